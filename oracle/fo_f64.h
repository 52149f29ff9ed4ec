// MinimalPathTracer evaluated in double precision (test infrastructure only): the FP64 CPU
// reference that converged images are compared against. It follows the same reference code as
// the FP32 restatement (MinimalPathTracer.rt.slang:294-468 tracePath / handleHit /
// generateScatterRay, StandardMaterial.slang:50-126, StandardBSDF.slang:125-313,
// PBRTDiffuseMaterialInstance.slang:54-78, ShadingUtils.slang:107-240) with every quantity after
// the scene description (camera basis, vertices, material records and texels, all FP32 inputs)
// in double: traversal, shading data, BSDF evaluation and sampling, throughput and the
// per-pixel mean. The uniform random numbers are the generator's own (24-bit quantised floats,
// SampleGeneratorInterface.slang:50-55), used as doubles. Emissive NEE (an extension) and the
// environment map are not part of this reference.
#pragma once
#include "fo_scene.h"

namespace fo {
namespace f64 {

using V3d = V3<double>;

inline V3d d3(V3f v) { return V3d(double(v.x), double(v.y), double(v.z)); }

struct HitD {
    uint32_t tri = kInvalid;
    double t = 0.0, u = 0.0, v = 0.0;
    bool valid() const { return tri != kInvalid; }
};

// Closest hit in double: the oracle's BVH2 (double slab tests on conservatively grown boxes) and
// the watertight triangle test (IntersectionHelpers.slang:118-187) in double; ties on t go to the
// smaller triangle id, as in the FP32 oracle.
inline HitD closestD(const Scene& s, V3d o, V3d d, double tmin, double tmax) {
    HitD best;
    auto test = [&](uint32_t tri) {
        V3d v[3];
        for (int j = 0; j < 3; j++) v[j] = d3(s.pos[s.idx[3 * tri + j]]);
        V3d b;
        double t;
        if (!intersectRayTriangle(o, d, v, b, t)) return;
        if (!(t >= tmin && t <= tmax)) return;
        if (best.valid() && !(t < best.t || (t == best.t && tri < best.tri))) return;
        best.tri = tri; best.t = t; best.u = b.y; best.v = b.z;
    };
    if (s.nodes.empty()) {
        for (uint32_t i = 0; i < s.nT; i++) test(i);
        return best;
    }
    double od[3] = {o.x, o.y, o.z};
    double inv[3] = {1.0 / d.x, 1.0 / d.y, 1.0 / d.z};
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const Scene::Node& n = s.nodes[stack[--sp]];
        double tE;
        if (!boxHit(n, od, inv, best.valid() ? best.t : tmax, tE)) continue;
        if (n.count > 0) {
            for (uint32_t i = 0; i < n.count; i++) test(s.order[n.first + i]);
        } else {
            stack[sp++] = n.right;
            stack[sp++] = n.left;
        }
    }
    return best;
}

struct ShadingD {
    V3d posW, V, faceN;
    double uvx, uvy;
    Frame<double> frame;
    bool frontFacing;
    uint32_t mat;
};
// Scene.slang:440-476 with world-space geometry.
inline ShadingD shadingD(const Scene& s, const HitD& h, V3d rayDir) {
    ShadingD sd;
    const uint32_t i0 = s.idx[3 * h.tri], i1 = s.idx[3 * h.tri + 1], i2 = s.idx[3 * h.tri + 2];
    const double b0 = 1.0 - h.u - h.v, b1 = h.u, b2 = h.v;
    const V3d p0 = d3(s.pos[i0]), p1 = d3(s.pos[i1]), p2 = d3(s.pos[i2]);
    sd.posW = p0 * b0 + p1 * b1 + p2 * b2;
    V3d n = d3(s.nrm[i0]) * b0 + d3(s.nrm[i1]) * b1 + d3(s.nrm[i2]) * b2;
    sd.uvx = s.uv[2 * i0] * b0 + s.uv[2 * i1] * b1 + s.uv[2 * i2] * b2;
    sd.uvy = s.uv[2 * i0 + 1] * b0 + s.uv[2 * i1 + 1] * b1 + s.uv[2 * i2 + 1] * b2;
    sd.faceN = normalize(cross(p1 - p0, p2 - p0));
    sd.V = -rayDir;
    sd.frame = frameFromNormal(normalize(n));
    sd.frontFacing = dot(sd.V, sd.faceN) >= 0.0;
    sd.mat = s.triMat[h.tri];
    return sd;
}
inline V3d rayOriginD(const ShadingD& sd, bool viewside = true) {
    return computeRayOrigin(sd.posW, (sd.frontFacing == viewside) ? sd.faceN : -sd.faceN);
}

// Bilinear fetch at mip 0 with wrap addressing, weights and sum in double.
inline V3d textureD(const Scene& s, int tex, double u, double v) {
    const int32_t* td = &s.texDesc[4 * tex];
    const int W = td[1], H = td[2];
    const double x = (u - std::floor(u)) * W - 0.5, y = (v - std::floor(v)) * H - 0.5;
    const double fx0 = std::floor(x), fy0 = std::floor(y), fx = x - fx0, fy = y - fy0;
    const int x0 = int(fx0), y0 = int(fy0);
    const int xa = ((x0 % W) + W) % W, xb = (((x0 + 1) % W) + W) % W;
    const int ya = ((y0 % H) + H) % H, yb = (((y0 + 1) % H) + H) % H;
    const float* base = &s.texels[td[0]];
    auto T = [&](int xx, int yy) { const float* q = base + 3 * (size_t(yy) * W + xx); return V3d(q[0], q[1], q[2]); };
    return (T(xa, ya) * (1.0 - fx) + T(xb, ya) * fx) * (1.0 - fy) + (T(xa, yb) * (1.0 - fx) + T(xb, yb) * fx) * fy;
}

struct MatD {
    Frame<double> sf;
    BsdfData<double> d;
    V3d emission;
};
// StandardMaterial.slang:50-126 / PBRTDiffuseMaterialInstance.slang:40-61 in double.
inline MatD materialD(const Scene& s, const ShadingD& sd) {
    const float* m = s.mat(sd.mat);
    MatD mi;
    mi.d.type = int(m[M_TYPE]);
    V3d base(m[M_BASE], m[M_BASE + 1], m[M_BASE + 2]);
    if (int(m[M_TEX]) >= 0) base = textureD(s, int(m[M_TEX]), sd.uvx, sd.uvy);
    mi.sf = sd.frame;
    if (!sd.frontFacing && matDoubleSided(m)) mi.sf.N = -mi.sf.N;
    mi.emission = V3d(0.0);
    if (mi.d.type == MAT_PBRT_DIFFUSE) {
        mi.d.diffuse = base;
        mi.d.specular = V3d(0.0);
        mi.d.roughness = 0.0;
        mi.d.metallic = 0.0;
        return mi;
    }
    const double ior = m[M_IOR], f = (ior - 1.0) / (ior + 1.0), F0 = f * f, metal = m[M_METAL];
    mi.d.diffuse = base + (V3d(0.0) - base) * metal;
    mi.d.specular = V3d(F0) + (base - V3d(F0)) * metal;
    mi.d.roughness = m[M_ROUGH];
    mi.d.metallic = metal;
    mi.d.eta = sd.frontFacing ? 1.0 / ior : ior;
    mi.d.specTrans = m[M_SPEC_TRANS];
    mi.d.diffTrans = m[M_DIFF_TRANS];
    mi.d.transmission = V3d(0.0);
    if (mi.d.specTrans > 0.0 || mi.d.diffTrans > 0.0) {
        const V3d t(m[M_TRANS], m[M_TRANS + 1], m[M_TRANS + 2]);
        mi.d.transmission = m[M_THIN] != 0.f ? t : V3d(std::sqrt(t.x), std::sqrt(t.y), std::sqrt(t.z));
    }
    if (sd.frontFacing) {
        const int etex = int(m[M_EMIT_TEX]) - 1;
        const V3d c = etex >= 0 ? textureD(s, etex, sd.uvx, sd.uvy) : V3d(m[M_EMIT], m[M_EMIT + 1], m[M_EMIT + 2]);
        mi.emission = c * double(m[M_EMIT_FACTOR]);
    }
    return mi;
}

// ShadingUtils.slang:155-181 / :214-240 in double.
inline bool validD(const ShadingD& sd, V3d wiL, V3d woL, V3d wo, bool reflectionOnly) {
    const bool isT = wiL.z * woL.z < 0.0;
    if (reflectionOnly ? std::min(wiL.z, woL.z) < kMinCosTheta : std::min(wiL.z, std::fabs(woL.z)) < kMinCosTheta)
        return false;
    const bool wiTop = sd.frontFacing, woTop = dot(wo, sd.faceN) >= 0.0;
    if (reflectionOnly) return wiTop == woTop;
    return isT ? wiTop != woTop : wiTop == woTop;
}

struct SampleD { V3d wo, weight; bool transmission; };
// StandardMaterialInstance.slang:89-110 + sampleReference :202-231 / PBRTDiffuseMaterialInstance.slang:66-78.
inline bool sampleD(const ShadingD& sd, const MatD& mi, SampleGen& sg, bool importance, SampleD& out) {
    const V3d wiL = mi.sf.toLocal(sd.V);
    V3d woL;
    double pdf;
    if (mi.d.type == MAT_PBRT_DIFFUSE) {
        const double ux = sg.next1D(), uy = sg.next1D();
        woL = sample_cosine_hemisphere_concentric(ux, uy, pdf);
        out.weight = mi.d.diffuse;
        out.wo = mi.sf.fromLocal(woL);
        out.transmission = false;
        return validD(sd, wiL, woL, out.wo, true) && pdf != 0.0;
    }
    StdBsdf<double> b;
    if (!importance) {
        const BsdfData<double>& q = mi.d;
        const bool anyDiffuse = q.diffuse.x > 0.0 || q.diffuse.y > 0.0 || q.diffuse.z > 0.0;
        const bool transmissive = q.specTrans > 0.0 || (anyDiffuse && q.specTrans < 1.0 && q.diffTrans > 0.0);
        const double ux = sg.next1D(), uy = sg.next1D();
        woL = sample_cosine_hemisphere_concentric(ux, uy, pdf);
        if (transmissive) {
            if (sg.next1D() < 0.5f) woL.z = -woL.z;
            pdf *= 0.5;
            if (std::min(std::fabs(wiL.z), std::fabs(woL.z)) < kMinCosTheta || pdf == 0.0) return false;
        } else if (std::min(wiL.z, woL.z) < kMinCosTheta || pdf == 0.0) {
            return false;
        }
        b.init(wiL, mi.d);
        out.weight = b.eval(wiL, woL) / pdf;
    } else {
        b.init(wiL, mi.d);
        bool delta;
        if (!b.sample(wiL, woL, pdf, out.weight, delta, sg)) return false;
    }
    out.wo = mi.sf.fromLocal(woL);
    out.transmission = woL.z < 0.0;
    return validD(sd, wiL, woL, out.wo, false);
}

// MinimalPathTracer.rt.slang:392-468 (no analytic lights, no environment map), one sample.
inline V3d tracePathD(const Scene& s, uint32_t maxBounces, bool computeDirect, bool importance, uint32_t px, uint32_t py,
                      uint32_t W, uint32_t H, uint32_t sampleNumber) {
    const Camera& c = s.cam;
    const double p_x = (px + 0.5) / W, p_y = (py + 0.5) / H;
    const double ndc_x = 2.0 * p_x - 1.0, ndc_y = -2.0 * p_y + 1.0;
    const V3d d = normalize(d3(c.U) * ndc_x + d3(c.V) * ndc_y + d3(c.W));
    const double invCos = 1.0 / dot(normalize(d3(c.W)), d);
    HitD h = closestD(s, d3(c.pos), d, c.nearZ * invCos, c.farZ * invCos);
    if (!h.valid()) return V3d(0.0);
    ShadingD sd = shadingD(s, h, d);
    MatD mi = materialD(s, sd);
    SampleGen sg = makeSampleGen(px, py, sampleNumber);
    V3d out = computeDirect ? mi.emission : V3d(0.0);
    V3d radiance(0.0), thp(1.0), rO, rD;
    uint32_t pathLength = 0;
    auto scatter = [&](const ShadingD& x, const MatD& m) -> bool {
        SampleD smp;
        if (!sampleD(x, m, sg, importance, smp)) return false;
        rO = rayOriginD(x, !smp.transmission);  // MinimalPathTracer.rt.slang:303-307
        rD = smp.wo;
        thp = thp * smp.weight;
        return anyPositive(thp);
    };
    if (!scatter(sd, mi)) return out;
    for (uint32_t depth = 0; depth <= maxBounces; depth++) {
        h = closestD(s, rO, rD, 0.0, std::numeric_limits<double>::max());
        if (!h.valid()) break;
        ShadingD x = shadingD(s, h, rD);
        MatD m = materialD(s, x);
        if (computeDirect || pathLength > 0) radiance = radiance + thp * m.emission;
        if (pathLength >= maxBounces) break;
        if (!scatter(x, m)) break;
        pathLength++;
    }
    return out + radiance;
}

}  // namespace f64
}  // namespace fo
