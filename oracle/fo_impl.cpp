// Oracle implementation: BVH2 build, light collection, LightBVH build/sampling, WAR primal.
// TEST INFRASTRUCTURE ONLY (see fo_core.h).
#include "fo_scene.h"
#include <array>
#include <stdexcept>

namespace fo {

// ---------------------------------------------------------------- oracle BVH2
// Median split on the longest centroid axis, leaves of <= 4 triangles. Independent of the
// product's builder on purpose: closest-hit answers must agree whatever the tree shape.
void buildOracleBVH(Scene& s) {
    s.nodes.clear();
    s.order.resize(s.nT);
    std::iota(s.order.begin(), s.order.end(), 0u);
    if (s.nT == 0) return;
    std::vector<V3f> cen(s.nT);
    std::vector<std::array<float, 6>> box(s.nT);
    for (uint32_t t = 0; t < s.nT; t++) {
        V3f v[3];
        s.triVerts(t, v);
        std::array<float, 6> b;
        for (int a = 0; a < 3; a++) {
            b[a] = std::min({v[0][a], v[1][a], v[2][a]});
            b[3 + a] = std::max({v[0][a], v[1][a], v[2][a]});
        }
        box[t] = b;
        cen[t] = V3f((b[0] + b[3]) * 0.5f, (b[1] + b[4]) * 0.5f, (b[2] + b[5]) * 0.5f);
    }
    struct Item { uint32_t node, first, count; };
    std::vector<Item> work;
    s.nodes.push_back({});
    work.push_back({0, 0, s.nT});
    while (!work.empty()) {
        Item it = work.back();
        work.pop_back();
        Scene::Node n{};
        float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
        float clo[3] = {1e30f, 1e30f, 1e30f}, chi[3] = {-1e30f, -1e30f, -1e30f};
        for (uint32_t i = it.first; i < it.first + it.count; i++) {
            uint32_t t = s.order[i];
            for (int a = 0; a < 3; a++) {
                lo[a] = std::min(lo[a], box[t][a]); hi[a] = std::max(hi[a], box[t][3 + a]);
                clo[a] = std::min(clo[a], cen[t][a]); chi[a] = std::max(chi[a], cen[t][a]);
            }
        }
        for (int a = 0; a < 3; a++) { n.lo[a] = lo[a]; n.hi[a] = hi[a]; }
        if (it.count <= 4) {
            n.first = it.first; n.count = it.count; n.left = n.right = 0;
            s.nodes[it.node] = n;
            continue;
        }
        int axis = 0;
        for (int a = 1; a < 3; a++) if (chi[a] - clo[a] > chi[axis] - clo[axis]) axis = a;
        uint32_t mid = it.first + it.count / 2;
        std::nth_element(s.order.begin() + it.first, s.order.begin() + mid, s.order.begin() + it.first + it.count,
                         [&](uint32_t x, uint32_t y) { return cen[x][axis] < cen[y][axis] || (cen[x][axis] == cen[y][axis] && x < y); });
        n.count = 0;
        n.left = uint32_t(s.nodes.size());
        s.nodes.push_back({});
        n.right = uint32_t(s.nodes.size());
        s.nodes.push_back({});
        s.nodes[it.node] = n;
        work.push_back({n.left, it.first, mid - it.first});
        work.push_back({n.right, mid, it.first + it.count - mid});
    }
}

// ---------------------------------------------------------------- light collection
// LightCollection.cpp:220-262 / UpdateTriangleVertices.cs.slang (world-space triangle data,
// Scene.slang:392-407 for normal and area) / FinalizeIntegration.cs.slang:103-111 (flux for a
// constant emissive colour). Triangles are listed in ascending global triangle order.
// Texture-space coverage of one texel: the polygon triangle-inside-square obtained by clipping
// against x >= i, x <= i + 1, y >= j, y <= j + 1 in turn, then its area.
static double texelCoverage(const std::vector<std::array<double, 2>>& tri, double i, double j) {
    std::vector<std::array<double, 2>> poly = tri, next;
    const double bounds[4] = {i, i + 1.0, j, j + 1.0};
    for (int e = 0; e < 4 && !poly.empty(); e++) {
        const int ax = e / 2;
        const bool keepAbove = (e % 2) == 0;
        next.clear();
        for (size_t k = 0; k < poly.size(); k++) {
            const auto& P = poly[k];
            const auto& Q = poly[(k + 1) % poly.size()];
            const double dp = keepAbove ? P[ax] - bounds[e] : bounds[e] - P[ax];
            const double dq = keepAbove ? Q[ax] - bounds[e] : bounds[e] - Q[ax];
            if (dp >= 0) next.push_back(P);
            if ((dp >= 0) != (dq >= 0)) {
                const double t = dp / (dp - dq);
                next.push_back({P[0] + t * (Q[0] - P[0]), P[1] + t * (Q[1] - P[1])});
            }
        }
        poly.swap(next);
    }
    double twice = 0.0;
    for (size_t k = 0; k < poly.size(); k++) {
        const auto& P = poly[k];
        const auto& Q = poly[(k + 1) % poly.size()];
        twice += P[0] * Q[1] - Q[0] * P[1];
    }
    return std::fabs(twice) * 0.5;
}

// EmissiveIntegrator.3d.slang + FinalizeIntegration.cs.slang:63-111: the coverage-weighted average of
// the nearest texels (wrap addressing) over the triangle's texture footprint, placed at
// (uv - floor(min uv)) * (W, H); a triangle with no coverage averages the texels at its vertices.
static V3f textureAverage(const Scene& s, int tex, const float uv[3][2]) {
    const int32_t* td = &s.texDesc[4 * tex];
    const long W = td[1], H = td[2];
    auto fetch = [&](long x, long y) {
        x = ((x % W) + W) % W;
        y = ((y % H) + H) % H;
        const float* q = &s.texels[size_t(td[0]) + 3 * size_t(y * W + x)];
        return V3f(q[0], q[1], q[2]);
    };
    const float ou = std::floor(std::min({uv[0][0], uv[1][0], uv[2][0]}));
    const float ov = std::floor(std::min({uv[0][1], uv[1][1], uv[2][1]}));
    std::vector<std::array<double, 2>> tri(3);
    for (int k = 0; k < 3; k++) tri[k] = {double((uv[k][0] - ou) * float(W)), double((uv[k][1] - ov) * float(H))};
    const double x0 = std::min({tri[0][0], tri[1][0], tri[2][0]}), x1 = std::max({tri[0][0], tri[1][0], tri[2][0]});
    const double y0 = std::min({tri[0][1], tri[1][1], tri[2][1]}), y1 = std::max({tri[0][1], tri[1][1], tri[2][1]});
    double acc[3] = {0.0, 0.0, 0.0}, weight = 0.0;
    for (long y = long(std::floor(y0)); y < long(std::ceil(y1)); y++)
        for (long x = long(std::floor(x0)); x < long(std::ceil(x1)); x++) {
            double w = texelCoverage(tri, double(x), double(y));
            if (w > 1.0) w = 1.0;
            if (!(w > 0.0)) continue;
            const V3f c = fetch(x, y);
            acc[0] += w * c.x; acc[1] += w * c.y; acc[2] += w * c.z;
            weight += w;
        }
    if (weight > 0.0) return V3f(float(acc[0] / weight), float(acc[1] / weight), float(acc[2] / weight));
    V3f sum(0.f);
    for (int k = 0; k < 3; k++)
        sum = sum + fetch(long(wrapCoord(uv[k][0]) * float(W)), long(wrapCoord(uv[k][1]) * float(H)));
    return sum / 3.f;
}

void buildLightCollection(Scene& s) {
    s.lights.clear();
    s.activeLights.clear();
    s.triToLight.assign(s.nT, kInvalid);
    for (uint32_t t = 0; t < s.nT; t++) {
        const float* m = s.mat(s.triMat[t]);
        if (int(m[M_TYPE]) != MAT_STANDARD) continue;
        const int etex = int(m[M_EMIT_TEX]) - 1;
        V3f Le = V3f(m[M_EMIT], m[M_EMIT + 1], m[M_EMIT + 2]) * m[M_EMIT_FACTOR];
        if (etex < 0 && !(Le.x > 0.f || Le.y > 0.f || Le.z > 0.f)) continue;
        if (etex >= 0 && !(m[M_EMIT_FACTOR] > 0.f)) continue;
        EmissiveTri e;
        s.triVerts(t, e.p);
        for (int j = 0; j < 3; j++) {
            e.uv[j][0] = s.uv[2 * s.idx[3 * t + j]];
            e.uv[j][1] = s.uv[2 * s.idx[3 * t + j] + 1];
        }
        V3f N = cross(e.p[1] - e.p[0], e.p[2] - e.p[0]);
        e.area = 0.5f * length(N);
        e.normal = normalize(N);
        e.Le = Le;
        if (etex >= 0) {
            Le = textureAverage(s, etex, e.uv) * m[M_EMIT_FACTOR];
            e.Le = V3f(m[M_EMIT_FACTOR]);
            e.etex = etex;
        }
        e.flux = luminance(Le) * e.area * Consts<float>::pi;
        e.tri = t;
        s.triToLight[t] = uint32_t(s.lights.size());
        s.lights.push_back(e);
    }
    // LightCollection.cpp:494-548: active list = lights with positive flux.
    for (uint32_t i = 0; i < s.lights.size(); i++)
        if (s.lights[i].flux > 0.f) s.activeLights.push_back(i);
}

// ---------------------------------------------------------------- LightBVH build
// LightBVHBuilder.cpp:35-882 (binned SAOH, 16 bins, <= 10 lights per leaf, leaves ASAP,
// flux pre-integration, lighting cones). Node partition uses a deterministic sort by
// (centroid, light index) in place of std::nth_element.
namespace {
constexpr float kInvalidCos = -1.f;  // LightBVHTypes.slang:43 (kInvalidCosConeAngle)
struct AABB {
    V3f lo{1e30f, 1e30f, 1e30f}, hi{-1e30f, -1e30f, -1e30f};
    bool valid() const { return lo.x <= hi.x && lo.y <= hi.y && lo.z <= hi.z; }
    void add(const AABB& o) {
        lo = V3f(std::min(lo.x, o.lo.x), std::min(lo.y, o.lo.y), std::min(lo.z, o.lo.z));
        hi = V3f(std::max(hi.x, o.hi.x), std::max(hi.y, o.hi.y), std::max(hi.z, o.hi.z));
    }
    void add(V3f p) { AABB b; b.lo = p; b.hi = p; add(b); }
    V3f center() const { return (lo + hi) * 0.5f; }
    V3f extent() const { return hi - lo; }
    float area() const { V3f e = extent(); return (e.x * e.y + e.x * e.z + e.y * e.z) * 2.f; }
};
inline float safeACos(float v) { return std::acos(std::clamp(v, -1.0f, 1.0f)); }
inline float sinFromCos(float c) { return std::sqrt(std::max(0.f, 1.f - c * c)); }
// LightBVHBuilder.cpp:67-88
float computeCosConeAngle(V3f coneDir, float cosTheta, V3f otherDir, float cosOther) {
    float cosResult = kInvalidCos;
    if (cosTheta != kInvalidCos && cosOther != kInvalidCos) {
        float cosDiff = dot(coneDir, otherDir);
        float sinDiff = sinFromCos(cosDiff);
        float sinOther = sinFromCos(cosOther);
        float cosTotal = cosOther * cosDiff - sinOther * sinDiff;
        float sinTotal = sinOther * cosDiff + cosOther * sinDiff;
        if (sinTotal > 0.f) cosResult = std::min(cosTheta, cosTotal);
    }
    return cosResult;
}
// LightBVHBuilder.cpp:94-109
V3f coneUnionOld(V3f aDir, float aCos, V3f bDir, float bCos, float& cosResult) {
    V3f dir = aDir + bDir;
    if (aCos == kInvalidCos || bCos == kInvalidCos || (dir.x == 0.f && dir.y == 0.f && dir.z == 0.f)) {
        cosResult = kInvalidCos;
        return V3f(0.f);
    }
    dir = normalize(dir);
    const float aDiff = safeACos(dot(dir, aDir));
    const float bDiff = safeACos(dot(dir, bDir));
    cosResult = std::cos(std::max(aDiff + std::acos(aCos), bDiff + std::acos(bCos)));
    return dir;
}
// LightBVHBuilder.cpp:659-665
float orientationCost(float theta_o) {
    float theta_w = std::min(theta_o + Consts<float>::pi_2, Consts<float>::pi);
    float s = std::sin(theta_o), c = std::cos(theta_o);
    return Consts<float>::two_pi * (1.0f - c) +
           Consts<float>::pi_2 * (2.0f * theta_w * s - std::cos(theta_o - 2.0f * theta_w) - 2.0f * theta_o * s + c);
}
// LightBVHBuilder.cpp:671-680 (surface area, pre-integration, lighting cones)
float evalSAOH(const AABB& b, float flux, float cosTheta) {
    float aabbCost = b.valid() ? b.area() : 0.f;
    float theta = cosTheta != kInvalidCos ? safeACos(cosTheta) : Consts<float>::pi;
    return flux * aabbCost * orientationCost(theta);
}
struct LTri { AABB bounds; V3f cone; float cosCone; float flux; uint32_t light; float c[3]; };
struct Builder {
    Scene& s;
    std::vector<LTri> tris;
    uint32_t maxLeaf = 10, binCount = 16;
    float currentNodeFlux = 0.f;

    // LightBVHBuilder.cpp:475-498
    V3f lightingCone(uint32_t b, uint32_t e, float& cosTheta) {
        V3f dir(0.f), sum(0.f);
        cosTheta = kInvalidCos;
        for (uint32_t i = b; i < e; i++) sum = sum + tris[i].cone;
        if (length(sum) >= std::numeric_limits<float>::min()) {
            dir = normalize(sum);
            cosTheta = 1.f;
            for (uint32_t i = b; i < e; i++) cosTheta = computeCosConeAngle(dir, cosTheta, tris[i].cone, tris[i].cosCone);
        }
        return dir;
    }
    // LightBVHBuilder.cpp:682-866; returns split index or 0 for "make a leaf".
    bool split(uint32_t b, uint32_t e, const AABB& nb, uint32_t& axisOut, uint32_t& idxOut) {
        float bestCost = std::numeric_limits<float>::infinity();
        bool found = false;
        V3f dims = nb.extent();
        uint32_t largest = dims.z >= dims.x && dims.z >= dims.y ? 2 : (dims.y >= dims.x && dims.y >= dims.z ? 1 : 0);
        struct Bin { AABB bounds; uint32_t count = 0; float flux = 0.f; V3f cone{0.f, 0.f, 0.f}; float cosCone = 1.f; };
        for (uint32_t dim = 0; dim < 3; dim++) {
            std::vector<Bin> bins(binCount);
            std::vector<float> costs(binCount - 1);
            float bmin = nb.lo[dim], bmax = nb.hi[dim];
            float w = bmax - bmin;
            float scale = w > std::numeric_limits<float>::min() ? float(binCount) / w : 0.f;
            auto binId = [&](const LTri& t) { float p = t.c[dim]; return std::min(uint32_t((p - bmin) * scale), binCount - 1); };
            for (uint32_t i = b; i < e; i++) {
                Bin& bn = bins[binId(tris[i])];
                bn.bounds.add(tris[i].bounds); bn.count += 1; bn.flux += tris[i].flux; bn.cone = bn.cone + tris[i].cone;
            }
            for (Bin& bn : bins) {
                bn.cosCone = length(bn.cone) < std::numeric_limits<float>::min() ? kInvalidCos : 1.0f;
                bn.cone = normalize(bn.cone);
            }
            for (uint32_t i = b; i < e; i++) {
                Bin& bn = bins[binId(tris[i])];
                bn.cosCone = computeCosConeAngle(bn.cone, bn.cosCone, tris[i].cone, tris[i].cosCone);
            }
            auto sweepCost = [&](size_t from, size_t to) {  // inclusive bin range
                Bin total;
                for (size_t j = from; j <= to; j++) {
                    total.bounds.add(bins[j].bounds); total.count += bins[j].count; total.flux += bins[j].flux;
                    total.cone = total.cone + bins[j].cone;
                }
                float cosTheta = kInvalidCos;
                if (length(total.cone) >= std::numeric_limits<float>::min()) {
                    cosTheta = 1.f;
                    V3f cd = normalize(total.cone);
                    for (size_t j = from; j <= to; j++) cosTheta = computeCosConeAngle(cd, cosTheta, bins[j].cone, bins[j].cosCone);
                }
                return evalSAOH(total.bounds, total.flux, cosTheta);
            };
            // Left sweep accumulates bins in order 0..i, right sweep i+1..n-1 (from the right);
            // the reference accumulates the running totals incrementally in those orders.
            {
                Bin total;
                for (size_t i = 0; i < costs.size(); i++) {
                    total.bounds.add(bins[i].bounds); total.count += bins[i].count; total.flux += bins[i].flux;
                    total.cone = total.cone + bins[i].cone;
                    float cosTheta = kInvalidCos;
                    if (length(total.cone) >= std::numeric_limits<float>::min()) {
                        cosTheta = 1.f;
                        V3f cd = normalize(total.cone);
                        for (size_t j = 0; j <= i; j++) cosTheta = computeCosConeAngle(cd, cosTheta, bins[j].cone, bins[j].cosCone);
                    }
                    costs[i] = evalSAOH(total.bounds, total.flux, cosTheta);
                }
                Bin tr;
                for (size_t i = costs.size(); i > 0; i--) {
                    tr.bounds.add(bins[i].bounds); tr.count += bins[i].count; tr.flux += bins[i].flux;
                    tr.cone = tr.cone + bins[i].cone;
                    float cosTheta = kInvalidCos;
                    if (length(tr.cone) >= std::numeric_limits<float>::min()) {
                        cosTheta = 1.f;
                        V3f cd = normalize(tr.cone);
                        for (size_t j = i; j <= costs.size(); j++) cosTheta = computeCosConeAngle(cd, cosTheta, bins[j].cone, bins[j].cosCone);
                    }
                    costs[i - 1] += evalSAOH(tr.bounds, tr.flux, cosTheta);
                }
            }
            (void)sweepCost;
            float axisBest = std::numeric_limits<float>::infinity();
            uint32_t axisIdx = b;
            for (uint32_t i = 0, triIdx = b; i < costs.size(); i++) {
                triIdx += bins[i].count;
                if (costs[i] < axisBest) { axisBest = costs[i]; axisIdx = triIdx; }
            }
            axisBest *= dims[largest] / dims[dim];
            if (axisIdx == b || axisIdx == e) continue;
            if (axisBest < bestCost) { bestCost = axisBest; axisOut = dim; idxOut = axisIdx; found = true; }
        }
        if (!found) {
            if (e - b <= maxLeaf) return false;
            // LightBVHBuilder.cpp:500-513 (computeSplitWithEqual fallback).
            axisOut = dims.z >= dims.x && dims.z >= dims.y ? 2 : (dims.y >= dims.x ? 1 : 0);
            idxOut = (b + e) / 2;
            return true;
        }
        if (e - b <= maxLeaf) {
            float cosTheta;
            lightingCone(b, e, cosTheta);
            float leafCost = evalSAOH(nb, currentNodeFlux, cosTheta);
            if (leafCost <= bestCost) return false;
        }
        return true;
    }
    // LightBVHBuilder.cpp:352-438
    uint32_t build(uint64_t bitmask, uint32_t depth, uint32_t b, uint32_t e) {
        float nodeFlux = 0.f;
        AABB nb;
        for (uint32_t i = b; i < e; i++) { nb.add(tris[i].bounds); nodeFlux += tris[i].flux; }
        currentNodeFlux = nodeFlux;
        uint32_t axis = 0, idx = 0;
        bool doSplit = (e - b) > maxLeaf && split(b, e, nb, axis, idx);
        uint32_t nodeIndex = uint32_t(s.lnodes.size());
        s.lnodes.push_back({});
        Scene::LNode n{};
        n.origin = nb.center();
        n.extent = nb.extent() * 0.5f;
        n.flux = nodeFlux;
        if (doSplit) {
            std::sort(tris.begin() + b, tris.begin() + e, [axis](const LTri& x, const LTri& y) {
                return x.c[axis] < y.c[axis] || (x.c[axis] == y.c[axis] && x.light < y.light);
            });
            if (depth >= 64) throw std::runtime_error("light BVH depth limit");
            uint32_t l = build(bitmask | (0ull << depth), depth + 1, b, idx);
            (void)l;
            uint32_t r = build(bitmask | (1ull << depth), depth + 1, idx, e);
            n.leaf = false;
            n.right = r;
            s.lnodes[nodeIndex] = n;
        } else {
            n.leaf = true;
            float c;
            n.coneDir = lightingCone(b, e, c);
            n.cosCone = c;
            n.triCount = e - b;
            n.triOffset = uint32_t(s.lTriIdx.size());
            for (uint32_t i = b; i < e; i++) {
                s.lTriIdx.push_back(tris[i].light);
                s.lBitmask[tris[i].light] = bitmask;
            }
            s.lnodes[nodeIndex] = n;
        }
        return nodeIndex;
    }
    // LightBVHBuilder.cpp:440-473
    V3f cones(uint32_t ni, float& cosCone) {
        Scene::LNode& n = s.lnodes[ni];
        if (!n.leaf) {
            float lc, rc;
            V3f ld = cones(ni + 1, lc);
            V3f rd = cones(n.right, rc);
            V3f d = coneUnionOld(ld, lc, rd, rc, cosCone);
            s.lnodes[ni].cosCone = cosCone;
            s.lnodes[ni].coneDir = d;
            return d;
        }
        cosCone = n.cosCone;
        return n.coneDir;
    }
};
}  // namespace

// ---------------------------------------------------------------- packed node precision
// The device light BVH stores PackedNode records (LightBVHTypes.slang:97-198): half-float box
// extents, the cone's cosine quantised to 16 bits and the cone axis as two 16-bit snorms in the
// octahedral map, while the origin and flux stay fp32. Sampling sees the decoded values, so the
// oracle keeps every node at the precision of its packed record.

// v rounded to the nearest IEEE binary16 value (ties to even; beyond 65504 to infinity), computed
// on the value in double precision rather than on bit fields.
float halfPrecision(float v) {
    if (!std::isfinite(v) || v == 0.f) return v;
    const double a = std::fabs(double(v));
    int e;
    std::frexp(a, &e);                            // a in [2^(e-1), 2^e)
    const double q = std::ldexp(1.0, std::max(e - 11, -24));  // spacing: 11 significant bits, or 2^-24
    const double r = a / q, fl = std::floor(r), frac = r - fl;
    const double n = frac > 0.5 ? fl + 1.0 : (frac < 0.5 ? fl : (std::fmod(fl, 2.0) == 0.0 ? fl : fl + 1.0));
    double res = n * q;
    if (res > 65504.0) res = std::numeric_limits<double>::infinity();
    return float(v < 0.f ? -res : res);
}
// floatToSnorm16 then unpackSnorm16 (FormatConversion.slang:235-251).
static float snorm16Precision(float v) {
    v = std::isnan(v) ? 0.f : std::min(std::max(v, -1.f), 1.f);
    const int bits = int(std::trunc(v * 32767.f + (v >= 0.f ? 0.5f : -0.5f)));
    return std::max(float(bits) / 32767.f, -1.0f);
}
// encodeNormal2x16 then decodeNormal2x16 (PackedFormats.slang:53-66, MathHelpers.slang:156-194).
V3f octNormalPrecision(V3f n) {
    auto wrap = [](float x, float y, float& ox, float& oy) {
        ox = (1.f - std::fabs(y)) * (x >= 0.f ? 1.f : -1.f);
        oy = (1.f - std::fabs(x)) * (y >= 0.f ? 1.f : -1.f);
    };
    const float inv = 1.f / (std::fabs(n.x) + std::fabs(n.y) + std::fabs(n.z));
    float px = n.x * inv, py = n.y * inv;
    if (n.z < 0.f) wrap(px, py, px, py);
    px = snorm16Precision(px);
    py = snorm16Precision(py);
    V3f d(px, py, 1.f - std::fabs(px) - std::fabs(py));
    if (d.z < 0.f) wrap(d.x, d.y, d.x, d.y);
    return normalize(d);
}
// (uint)((cos + 1) * 32767), decoded as packed / 32767 - 1.
float conePrecision(float c) { return float(uint32_t((c + 1.f) * 32767.f)) * (1.f / 32767.f) - 1.f; }
static void toPackedPrecision(Scene::LNode& n) {
    n.extent = V3f(halfPrecision(n.extent.x), halfPrecision(n.extent.y), halfPrecision(n.extent.z));
    n.cosCone = conePrecision(n.cosCone);
    n.coneDir = octNormalPrecision(n.coneDir);
}

void buildLightBVH(Scene& s) {
    s.lnodes.clear();
    s.lTriIdx.clear();
    s.lBitmask.assign(s.lights.size(), ~0ull);
    Builder B{s, {}};
    for (uint32_t i = 0; i < s.lights.size(); i++) {
        const EmissiveTri& e = s.lights[i];
        if (!(e.flux > 0.f)) continue;
        LTri t;
        for (int j = 0; j < 3; j++) t.bounds.add(e.p[j]);
        V3f c = t.bounds.center();
        t.c[0] = c.x; t.c[1] = c.y; t.c[2] = c.z;
        t.cone = e.normal;
        t.cosCone = 1.f;
        t.flux = e.flux;
        t.light = i;
        B.tris.push_back(t);
    }
    if (B.tris.empty()) return;
    B.build(0ull, 0, 0, uint32_t(B.tris.size()));
    float c;
    B.cones(0, c);
    // The builder works in fp32; the records it writes hold the packed precision.
    for (Scene::LNode& n : s.lnodes) toPackedPrecision(n);
}

// ---------------------------------------------------------------- animation update
// UpdateTriangleVertices.cs.slang:42-63 (LightCollection::updateTrianglePositions,
// LightCollection.cpp:550-574): every emissive triangle's world-space vertices, then
// Scene::computeFaceNormalAndAreaW (Scene.slang:392-407); flux is left as built.
void updateLightPositions(Scene& s) {
    for (EmissiveTri& e : s.lights) {
        s.triVerts(e.tri, e.p);
        V3f e0 = e.p[1] - e.p[0], e1 = e.p[2] - e.p[0];
        V3f N = cross(e0, e1);
        e.area = 0.5f * length(N);
        e.normal = normalize(N);
    }
}

namespace {
// LightBVHTypes.slang:56-69 (AABB stored as origin and half extent).
void setNodeAABB(Scene::LNode& n, V3f lo, V3f hi) {
    n.origin = (hi + lo) * 0.5f;
    n.extent = (hi - lo) * 0.5f;
}
float sinOfCosine(float c) { return std::sqrt(std::max(0.0f, 1.0f - c * c)); }
}  // namespace

// LightBVH::refit (LightBVH.cpp:49-89) with LightBVHRefit.cs.slang: all leaves (updateLeafNodes,
// :46-93), then the internal nodes level by level from the deepest (updateInternalNodes,
// :101-170). Node depths follow LightBVH::updateNodeIndices (LightBVH.cpp:219-250).
void refitLightBVH(Scene& s) {
    const uint32_t n = uint32_t(s.lnodes.size());
    if (n == 0) return;
    std::vector<uint32_t> depth(n, 0);
    uint32_t height = 0;
    for (uint32_t i = 0; i < n; i++) {
        if (s.lnodes[i].leaf) continue;
        depth[i + 1] = depth[i] + 1;
        depth[s.lnodes[i].right] = depth[i] + 1;
        height = std::max(height, depth[i] + 1);
    }
    const float kFltMaxF = std::numeric_limits<float>::max(), kFltMinF = std::numeric_limits<float>::min();
    for (uint32_t i = 0; i < n; i++) {
        Scene::LNode& node = s.lnodes[i];
        if (!node.leaf) continue;
        V3f lo(kFltMaxF), hi(-kFltMaxF), sum(0.f);
        for (uint32_t k = 0; k < node.triCount; k++) {
            const EmissiveTri& e = s.lights[s.lTriIdx[node.triOffset + k]];
            for (int j = 0; j < 3; j++) {
                lo = V3f(std::min(lo.x, e.p[j].x), std::min(lo.y, e.p[j].y), std::min(lo.z, e.p[j].z));
                hi = V3f(std::max(hi.x, e.p[j].x), std::max(hi.y, e.p[j].y), std::max(hi.z, e.p[j].z));
            }
            sum = sum + e.normal;
        }
        setNodeAABB(node, lo, hi);
        float len = length(sum);
        V3f dir = sum / len;
        float cosCone = -1.f;  // kInvalidCosConeAngle
        if (len >= kFltMinF) {
            cosCone = 1.f;
            for (uint32_t k = 0; k < node.triCount; k++)
                cosCone = std::min(cosCone, dot(dir, s.lights[s.lTriIdx[node.triOffset + k]].normal));
            cosCone = std::max(cosCone, -1.f);
        }
        node.cosCone = cosCone;
        node.coneDir = dir;
        toPackedPrecision(node);
    }
    for (int d = int(height) - 1; d >= 0; d--) {
        for (uint32_t i = 0; i < n; i++) {
            Scene::LNode& node = s.lnodes[i];
            if (node.leaf || depth[i] != uint32_t(d)) continue;
            const Scene::LNode& L = s.lnodes[i + 1];
            const Scene::LNode& R = s.lnodes[node.right];
            V3f lmin = L.origin - L.extent, lmax = L.origin + L.extent;
            V3f rmin = R.origin - R.extent, rmax = R.origin + R.extent;
            setNodeAABB(node, V3f(std::min(lmin.x, rmin.x), std::min(lmin.y, rmin.y), std::min(lmin.z, rmin.z)),
                        V3f(std::max(lmax.x, rmax.x), std::max(lmax.y, rmax.y), std::max(lmax.z, rmax.z)));
            V3f sum = L.coneDir + R.coneDir;
            float len = length(sum);
            V3f dir = sum / len;
            float cosCone = -1.f;
            if (len >= kFltMinF && L.cosCone != -1.f && R.cosCone != -1.f) {
                // Rotate each child's (cos, sin) of its offset from dir by its own spread.
                float cosLD = dot(dir, L.coneDir), sinLD = sinOfCosine(cosLD);
                float cosRD = dot(dir, R.coneDir), sinRD = sinOfCosine(cosRD);
                float sinLC = sinOfCosine(L.cosCone), sinRC = sinOfCosine(R.cosCone);
                float sinLT = sinLC * cosLD + sinLD * L.cosCone;
                float sinRT = sinRC * cosRD + sinRD * R.cosCone;
                if (sinLT > 0.f && sinRT > 0.f) {
                    float cosLT = L.cosCone * cosLD - sinLC * sinLD;
                    float cosRT = R.cosCone * cosRD - sinRC * sinRD;
                    cosCone = std::max(std::min(cosLT, cosRT), -1.f);
                }
            }
            node.cosCone = cosCone;
            node.coneDir = dir;
            toPackedPrecision(node);
        }
    }
}

// ---------------------------------------------------------------- LightBVH sampling
// LightBVHSampler.slang:153-170
static inline float cosSubClamped(float sinA, float cosA, float sinB, float cosB) {
    if (cosA > cosB) return 1.f;
    return cosA * cosB + sinA * sinB;
}
static inline float sinSubClamped(float sinA, float cosA, float sinB, float cosB) {
    if (cosA > cosB) return 0.f;
    return sinA * cosB - cosA * sinB;
}
// LightBVHSampler.slang:181-210 (Sphere bound) + GeometryHelpers.slang:222-240
static float boundCosineTerm(V3f posW, V3f n, V3f center, V3f extent, float& cosThetaCone) {
    float sinThetaCone = 0.f;
    float sqrRadius = dot(extent, extent);
    V3f cv = center - posW;
    float cd2 = dot(cv, cv);
    if (cd2 < sqrRadius) { sinThetaCone = 0.f; cosThetaCone = -1.f; }
    else {
        float sin2 = sqrRadius / cd2;
        cosThetaCone = std::sqrt(1.f - sin2);
        sinThetaCone = std::sqrt(sin2);
    }
    V3f L = normalize(center - posW);
    float cosThetaL = std::clamp(dot(n, L), -1.f, 1.f);
    float sinThetaL = std::sqrt(1.f - cosThetaL * cosThetaL);
    return saturate(cosSubClamped(sinThetaL, cosThetaL, sinThetaCone, cosThetaCone));
}
// LightBVHSampler.slang:219-273
float lbvhImportance(const Scene& s, V3f posW, V3f n, uint32_t node) {
    const Scene::LNode& a = s.lnodes[node];
    float flux = a.flux;
    float distance = length(a.origin - posW);
    float cosThetaBoundingCone = 0.f;
    float NdotL = boundCosineTerm(posW, n, a.origin, a.extent, cosThetaBoundingCone);
    float orientationWeight = 1.f;
    float cosConeAngle = a.cosCone;
    V3f dirToAabb = (a.origin - posW) / distance;
    if (cosConeAngle != kInvalidCos && cosConeAngle > 0.f) {
        float sinConeAngle = std::sqrt(std::max(0.f, 1.f - cosConeAngle * cosConeAngle));
        float cosTheta = dot(a.coneDir, -dirToAabb);
        float sinTheta = std::sqrt(std::max(0.f, 1.f - cosTheta * cosTheta));
        float sinThetaBC = std::sqrt(std::max(0.f, 1 - cosThetaBoundingCone * cosThetaBoundingCone));
        float cosTheta0 = cosSubClamped(sinTheta, cosTheta, sinConeAngle, cosConeAngle);
        float sinTheta0 = sinSubClamped(sinTheta, cosTheta, sinConeAngle, cosConeAngle);
        float cosThetaPrime = cosSubClamped(sinTheta0, cosTheta0, sinThetaBC, cosThetaBoundingCone);
        orientationWeight = std::max(0.f, cosThetaPrime);
    }
    float halfRadius = std::max(a.extent.x, std::max(a.extent.y, a.extent.z));
    distance = std::max(halfRadius, distance);
    return (flux * NdotL) * orientationWeight / (distance * distance);
}
// LightBVHSampler.slang:284-323 + :375-419 (uniform triangle pick) + :430-446
bool lbvhSample(const Scene& s, V3f posW, V3f n, float u, float& pdf, uint32_t& light) {
    if (s.lnodes.empty()) return false;
    pdf = 1.f;
    uint32_t node = 0;
    while (!s.lnodes[node].leaf) {
        uint32_t l = node + 1, r = s.lnodes[node].right;
        float li = lbvhImportance(s, posW, n, l), ri = lbvhImportance(s, posW, n, r);
        float total = li + ri;
        if (total == 0.f) return false;
        float pLeft = li / total, pRight = 1.0f - pLeft;
        if (u < pLeft) { u = u / pLeft; pdf *= pLeft; node = l; }
        else { u = (u - pLeft) / pRight; pdf *= pRight; node = r; }
    }
    const Scene::LNode& leaf = s.lnodes[node];
    uint32_t idx = std::min(uint32_t(u * float(leaf.triCount)), leaf.triCount - 1);
    light = s.lTriIdx[leaf.triOffset + idx];
    pdf *= 1.0f / float(leaf.triCount);
    return true;
}
// LightBVHSampler.slang:108-124 + :455-492 + :502-533
float lbvhSelectionPdf(const Scene& s, V3f posW, V3f n, uint32_t light) {
    if (s.lnodes.empty() || light >= s.lBitmask.size()) return 0.f;
    uint64_t bitmask = s.lBitmask[light];
    float traversalPdf = 1.f;
    uint32_t node = 0;
    while (!s.lnodes[node].leaf) {
        uint32_t l = node + 1, r = s.lnodes[node].right;
        float li = lbvhImportance(s, posW, n, l), ri = lbvhImportance(s, posW, n, r);
        float total = li + ri;
        if (total == 0.f) return 0.f;
        float pLeft = li / total, pRight = 1.0f - pLeft;
        if ((bitmask & 1ull) == 0) { traversalPdf *= pLeft; node = l; }
        else { traversalPdf *= pRight; node = r; }
        bitmask >>= 1;
    }
    if (traversalPdf == 0.f) return 0.f;
    return traversalPdf * (1.0f / float(s.lnodes[node].triCount));
}

// ---------------------------------------------------------------- WAR primal
// DiffSceneQuery.slang:131-177 (AttachToRay): Moeller distance, area barycentrics,
// interpolated (unnormalised) vertex normal.
void computeIntersectionAD(const Scene& s, const Hit& h, V3f o, V3f d, V3f& posW, V3f& bary, V3f& faceN, V3f& normalW) {
    V3f v[3];
    s.triVerts(h.tri, v);
    float t = computeRayDistanceToTriangle(v, o, d);
    V3f p = o + d * t;
    bary = computeBarycentrics(v, p);
    posW = v[0] * bary.x + v[1] * bary.y + v[2] * bary.z;
    faceN = normalize(cross(v[1] - v[0], v[2] - v[0]));
    uint32_t i0 = s.idx[3 * h.tri], i1 = s.idx[3 * h.tri + 1], i2 = s.idx[3 * h.tri + 2];
    normalW = s.nrm[i0] * bary.x + s.nrm[i1] * bary.y + s.nrm[i2] * bary.z;
}

// WARDiffPathTracer.rt.slang:60-63
static inline void sampleGaussian(float ux, float uy, float sigma, float& gx, float& gy) {
    float r = std::sqrt(-2.f * std::log(ux));
    gx = (r * std::cos(Consts<float>::two_pi * uy)) * sigma;
    gy = (r * std::sin(Consts<float>::two_pi * uy)) * sigma;
}

V3f warTracePaths(const Scene& s, const WarParams& P, uint32_t px, uint32_t py, uint32_t W, uint32_t H, uint32_t frameSeed,
                  Tracer& tr) {
    SampleGen sg = makeSampleGen(px, py, frameSeed / 2);
    for (int i = 0; i < 6; i++) sg.next1D();
    float pfx = float(px) + 0.5f, pfy = float(py) + 0.5f;
    float ux = sg.next1D(), uy = sg.next1D();
    float jx, jy;
    sampleGaussian(ux, uy, 0.5f, jx, jy);
    if (P.useAntithetic && frameSeed % 2 == 0) { jx = -jx; jy = -jy; }
    float sx = pfx + jx, sy = pfy + jy;
    // tracePath (WARDiffPathTracer.rt.slang:75-111), useWAR = false.
    V3f radiance(0.f), thp(1.f);
    uint32_t pathLen = 0;
    V3f prevNormal(0.f); float prevPdf = 0.f;
    // computeCameraRayDirection (DiffSceneQuery.slang:252-261)
    float p_x = sx / float(W), p_y = sy / float(H);
    float ndc_x = 2.f * p_x + -1.f, ndc_y = -2.f * p_y + 1.f;
    V3f dir = s.cam.U * ndc_x + s.cam.V * ndc_y + s.cam.W;
    float dl = length(dir);
    dir = dl > 0.f ? dir / dl : V3f(0.f);
    V3f rayO = s.cam.pos, rayD = dir;
    Hit h = tr.closest(rayO, rayD, 0.f, std::numeric_limits<float>::max());
    if (!h.valid()) return radiance;
    const float FLTMAX = std::numeric_limits<float>::max();
    for (;;) {
        V3f posW, bary, faceN, normalW;
        computeIntersectionAD(s, h, rayO, rayD, posW, bary, faceN, normalW);
        // PTUtils.slang:97-123: shading data from the triangle hit with barycentrics (b.y, b.z).
        Hit th = h; th.u = bary.y; th.v = bary.z;
        ShadingData sd = loadShadingData(s, th, rayD);
        sd.V = -rayD;
        sd.frame = frameFromNormal(normalW);
        MatInst mi = makeMatInst(s, sd);
        bool isPrimary = pathLen == 0;
        bool computeEmissive = isPrimary || (!P.useNEE || P.useMIS);
        if (computeEmissive && anyPositive(mi.emission)) {
            float misWeight = 1.f;
            if (!isPrimary && P.useNEE && P.useMIS) {
                float lp = lightPdf(s, P.lightSampler, rayO, prevNormal, s.triToLight[h.tri], sd.posW, orientedFaceN(sd));
                misWeight = evalMIS(1.f, prevPdf, 1.f, lp);
            }
            radiance = radiance + thp * mi.emission * misWeight;
        }
        if (pathLen >= P.maxBounces + 1) break;
        if (P.useNEE) {
            LightSample ls;
            if (sampleLight(s, P.lightSampler, posW, normalW, sg, ls)) {
                V3f Li = ls.pdf > 0.f ? ls.Le / ls.pdf : V3f(0.f);
                V3f lightPos = computeRayOrigin(ls.posW, ls.normalW);
                V3f origin = computeRayOrigin(posW, normalW);
                V3f toLight = lightPos - origin;
                float distance = length(toLight);
                V3f ldir = normalize(toLight);
                if (anyPositive(Li) && tr.visible(origin, ldir, distance)) {
                    // computeNEE (PTUtils.slang:211-249)
                    V3f wo = ls.posW - posW;
                    float dist = length(wo);
                    wo = wo / dist;
                    float a2s = (dist * dist) / dot(ls.normalW, -wo);
                    V3f f = miEval(sd, mi, wo);
                    V3f weight = Li * f * a2s / a2s;
                    if (P.useMIS) {
                        float scatterPdf = miEvalPdf(sd, mi, ldir);
                        if (!P.useBSDFSampling) {
                            V3f wiL = mi.sf.toLocal(sd.V), woL = mi.sf.toLocal(ldir);
                            scatterPdf = (mi.d.type == MAT_PBRT_DIFFUSE) ? scatterPdf
                                         : (validHemisphere(sd, wiL, woL, ldir) && std::min(wiL.z, woL.z) >= kMinCosTheta ? woL.z * Consts<float>::inv_pi : 0.f);
                        }
                        weight = weight * evalMIS(1.f, ls.pdf, 1.f, scatterPdf);
                    }
                    radiance = radiance + thp * weight;
                }
            }
        }
        // generateScatterRay (PTUtils.slang:136-172)
        BsdfSample bs;
        if (!(miSample(sd, mi, sg, bs, P.useBSDFSampling) && bs.pdf > 0.f)) break;
        rayD = bs.wo;
        rayO = computeRayOrigin(posW, sd.frontFacing ? sd.faceN : -sd.faceN);
        V3f f = miEval(sd, mi, rayD);
        thp = thp * (f / bs.pdf);
        prevNormal = orientedFaceN(sd);
        prevPdf = bs.pdf;
        if (!anyPositive(thp)) break;
        pathLen++;
        h = tr.closest(rayO, rayD, 0.f, FLTMAX);
        if (!h.valid()) break;
    }
    return radiance;
}

}  // namespace fo
