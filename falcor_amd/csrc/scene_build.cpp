// Host-side scene build for the fhp core (see scene_build.h).
#include "scene_build.h"
#include "fhp_scene.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <stdexcept>

namespace fhp {

// Camera.cpp:187-193 (right-handed look-at basis scaled by focal distance and FOV).
void computeCamera(HostScene& s) {
    const float* c = s.cam;
    f3 pos = mk3(c[0], c[1], c[2]), target = mk3(c[3], c[4], c[5]), up = mk3(c[6], c[7], c[8]);
    float fovY = c[9], aspect = c[10], focalDistance = c[11];
    f3 W = normalize(target - pos) * focalDistance;
    f3 U = normalize(cross(W, up));
    f3 V = normalize(cross(U, W));
    const float ulen = focalDistance * std::tan(fovY * 0.5f) * aspect;
    U = U * ulen;
    const float vlen = focalDistance * std::tan(fovY * 0.5f);
    V = V * vlen;
    s.camU[0] = U.x; s.camU[1] = U.y; s.camU[2] = U.z;
    s.camV[0] = V.x; s.camV[1] = V.y; s.camV[2] = V.z;
    s.camW[0] = W.x; s.camW[1] = W.y; s.camW[2] = W.z;

    // viewProj = perspective(fovY, aspect, near, far) * lookAt(pos, target, up), in float.
    const float nearZ = c[12], farZ = c[13];
    const f3 f = normalize(pos - target);
    const f3 r = normalize(cross(up, f));
    const f3 u = cross(f, r);
    const float view[16] = {r.x, r.y, r.z, -dot(r, pos), u.x, u.y, u.z, -dot(u, pos),
                            f.x, f.y, f.z, -dot(f, pos), 0.f, 0.f, 0.f, 1.f};
    const float t = std::tan(fovY / 2.f);
    float proj[16] = {};
    proj[0] = 1.f / (aspect * t);
    proj[5] = 1.f / t;
    proj[10] = farZ / (nearZ - farZ);
    proj[11] = -(farZ * nearZ) / (farZ - nearZ);
    proj[14] = -1.f;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            float acc = 0.f;
            for (int k = 0; k < 4; k++) acc += proj[4 * i + k] * view[4 * k + j];
            s.viewProj[4 * i + j] = acc;
        }
}

void computeHitInfo(HostScene& s, bool useCompression) {
    // allocateBits (HitInfo.cpp:44-49): bits to store values 0 .. count - 1.
    auto bits = [](uint64_t count) -> uint32_t {
        uint32_t b = 0;
        while (count > 1 && (uint64_t(1) << b) < count) b++;
        return b;
    };
    const uint32_t nMesh = uint32_t(s.meshOff.size() - 1);
    s.meshTriOff.assign(nMesh + 1, 0u);
    s.meshTriContiguous = true;
    std::vector<uint32_t> count(nMesh, 0u);
    uint32_t prev = 0;
    for (uint32_t t = 0; t < s.nT; t++) {
        const uint32_t v = s.idx[3 * size_t(t)];
        const uint32_t m = uint32_t(std::upper_bound(s.meshOff.begin(), s.meshOff.end(), v) - s.meshOff.begin()) - 1;
        if (m < prev) s.meshTriContiguous = false;
        prev = m;
        count[m]++;
    }
    uint32_t maxPrim = 0;
    for (uint32_t m = 0; m < nMesh; m++) {
        s.meshTriOff[m + 1] = s.meshTriOff[m] + count[m];
        maxPrim = std::max(maxPrim, count[m]);
    }
    // HitType::Count = 7 (None, Triangle, Volume, DisplacedTriangle, Curve, reserved, SDFGrid).
    s.hitTypeBits = bits(7);
    s.hitInstBits = bits(nMesh);
    s.hitPrimBits = bits(maxPrim);
    if (s.hitPrimBits > 32 || s.hitTypeBits + s.hitInstBits > 32)
        throw std::runtime_error("Scene requires > 64 bits for encoding hit info header.");
    const uint32_t kCompressedTypeBits = 2;
    s.hitCompressed = useCompression && kCompressedTypeBits + s.hitInstBits + s.hitPrimBits <= 32;
    if (s.hitCompressed) s.hitTypeBits = kCompressedTypeBits;
}

// ------------------------------------------------------------------ BVH4 build
// Binned SAH binary tree (32 bins, three axes), collapsed into a 4-wide tree by repeatedly
// opening the child with the largest surface area. Node boxes are padded outward so the slab
// test never culls a box that holds the watertight hit.
namespace {
struct Box {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const Box& b) {
        for (int a = 0; a < 3; a++) { lo[a] = std::min(lo[a], b.lo[a]); hi[a] = std::max(hi[a], b.hi[a]); }
    }
    void grow(const float* p) {
        for (int a = 0; a < 3; a++) { lo[a] = std::min(lo[a], p[a]); hi[a] = std::max(hi[a], p[a]); }
    }
    double area() const {
        if (lo[0] > hi[0]) return 0.0;
        double dx = double(hi[0]) - lo[0], dy = double(hi[1]) - lo[1], dz = double(hi[2]) - lo[2];
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    }
};
struct N2 {
    Box box;
    int32_t l = -1, r = -1;
    uint32_t first = 0, count = 0;
};
struct Builder2 {
    std::vector<Box> pb;
    std::vector<float> cen;  // 3 per prim
    std::vector<uint32_t> ord;
    std::vector<N2> nodes;
    uint32_t maxLeaf;

    int32_t build(uint32_t b, uint32_t e) {
        int32_t ni = int32_t(nodes.size());
        nodes.emplace_back();
        Box bb, cb;
        for (uint32_t i = b; i < e; i++) { bb.grow(pb[ord[i]]); cb.grow(&cen[3 * ord[i]]); }
        nodes[ni].box = bb;
        const uint32_t n = e - b;
        auto makeLeaf = [&]() { nodes[ni].first = b; nodes[ni].count = n; return ni; };
        if (n == 1) return makeLeaf();
        constexpr int kBins = 32;
        double best = std::numeric_limits<double>::infinity();
        int bestAxis = -1, bestBin = -1;
        const double pa = std::max(bb.area(), 1e-30);
        for (int a = 0; a < 3; a++) {
            float ext = cb.hi[a] - cb.lo[a];
            if (!(ext > 0.f)) continue;
            float scale = float(kBins) / ext;
            Box bins[kBins];
            uint32_t cnt[kBins] = {};
            for (uint32_t i = b; i < e; i++) {
                uint32_t p = ord[i];
                int k = std::min(kBins - 1, int((cen[3 * p + a] - cb.lo[a]) * scale));
                bins[k].grow(pb[p]);
                cnt[k]++;
            }
            double rightA[kBins];
            uint32_t rightN[kBins];
            Box acc;
            uint32_t an = 0;
            for (int k = kBins - 1; k > 0; k--) {
                acc.grow(bins[k]); an += cnt[k];
                rightA[k] = acc.area(); rightN[k] = an;
            }
            Box lacc;
            uint32_t ln = 0;
            for (int k = 0; k < kBins - 1; k++) {
                lacc.grow(bins[k]); ln += cnt[k];
                if (ln == 0 || rightN[k + 1] == 0) continue;
                double cost = 1.0 + (lacc.area() * ln + rightA[k + 1] * rightN[k + 1]) / pa;
                if (cost < best) { best = cost; bestAxis = a; bestBin = k; }
            }
        }
        if (n <= maxLeaf && (bestAxis < 0 || double(n) <= best)) return makeLeaf();
        uint32_t mid;
        if (bestAxis < 0) {
            mid = b + n / 2;  // all centroids coincide: split the list in half
        } else {
            float ext = cb.hi[bestAxis] - cb.lo[bestAxis];
            float scale = float(kBins) / ext;
            auto it = std::partition(ord.begin() + b, ord.begin() + e, [&](uint32_t p) {
                return std::min(kBins - 1, int((cen[3 * p + bestAxis] - cb.lo[bestAxis]) * scale)) <= bestBin;
            });
            mid = uint32_t(it - ord.begin());
            if (mid == b || mid == e) mid = b + n / 2;
        }
        int32_t l = build(b, mid);
        int32_t r = build(mid, e);
        nodes[ni].l = l;
        nodes[ni].r = r;
        return ni;
    }
};
}  // namespace

void buildBVH4(HostScene& s, uint32_t maxLeaf) {
    if (maxLeaf < 1 || maxLeaf > 8) throw std::runtime_error("BVH leaf size must be in [1, 8]");
    if (s.nT == 0) throw std::runtime_error("scene has no triangles");
    if (s.nT >= (1u << 28)) throw std::runtime_error("too many triangles for the BVH4 leaf encoding");
    Builder2 B;
    B.maxLeaf = maxLeaf;
    B.pb.resize(s.nT);
    B.cen.resize(3 * size_t(s.nT));
    B.ord.resize(s.nT);
    std::iota(B.ord.begin(), B.ord.end(), 0u);
    float sceneLo[3] = {INFINITY, INFINITY, INFINITY}, sceneHi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t t = 0; t < s.nT; t++) {
        Box b;
        for (int j = 0; j < 3; j++) b.grow(&s.pos[3 * size_t(s.idx[3 * size_t(t) + j])]);
        B.pb[t] = b;
        for (int a = 0; a < 3; a++) {
            B.cen[3 * size_t(t) + a] = 0.5f * (b.lo[a] + b.hi[a]);
            sceneLo[a] = std::min(sceneLo[a], b.lo[a]);
            sceneHi[a] = std::max(sceneHi[a], b.hi[a]);
        }
    }
    B.nodes.reserve(2 * size_t(s.nT));
    B.build(0, s.nT);

    // Triangles in leaf order: (v0, id) (v1, 0) (v2, 0).
    s.triPos.assign(12 * size_t(s.nT), 0.f);
    for (uint32_t i = 0; i < s.nT; i++) {
        uint32_t t = B.ord[i];
        for (int j = 0; j < 3; j++) {
            const float* p = &s.pos[3 * size_t(s.idx[3 * size_t(t) + j])];
            float* q = &s.triPos[12 * size_t(i) + 4 * j];
            q[0] = p[0]; q[1] = p[1]; q[2] = p[2];
        }
        uint32_t idbits = t;
        std::memcpy(&s.triPos[12 * size_t(i) + 3], &idbits, 4);
    }

    float ext = 0.f;
    for (int a = 0; a < 3; a++) ext = std::max(ext, std::max(std::fabs(sceneLo[a]), std::fabs(sceneHi[a])));
    const float absPad = ext * 1e-6f + 1e-30f;
    s.bvhAbsPad = absPad;
    auto padLo = [&](float v) { return v - (std::fabs(v) * 1e-6f + absPad); };
    auto padHi = [&](float v) { return v + (std::fabs(v) * 1e-6f + absPad); };

    // Collapse into BVH4 nodes, depth-first.
    std::vector<float>& out = s.bvh;
    out.clear();
    s.bvhDepth = 0;
    s.stackNeed = 0;
    s.bvhLeaves = 0;
    s.bvhSah = 0.0;
    const double rootArea = std::max(B.nodes[0].box.area(), 1e-30);
    struct Ret { uint32_t ref, depth, need; };
    std::function<Ret(int32_t, uint32_t)> collapse = [&](int32_t n2, uint32_t depth) -> Ret {
        const N2& nd = B.nodes[n2];
        if (nd.l < 0) {
            s.bvhLeaves++;
            s.bvhSah += nd.box.area() / rootArea * nd.count;
            return Ret{makeLeaf(nd.first, nd.count), depth, 0};
        }
        std::vector<int32_t> ch = {nd.l, nd.r};
        while (ch.size() < 4) {
            int pick = -1;
            double pa = -1.0;
            for (size_t i = 0; i < ch.size(); i++) {
                const N2& c = B.nodes[ch[i]];
                if (c.l >= 0 && c.box.area() > pa) { pa = c.box.area(); pick = int(i); }
            }
            if (pick < 0) break;
            int32_t opened = ch[pick];
            ch.erase(ch.begin() + pick);
            ch.push_back(B.nodes[opened].l);
            ch.push_back(B.nodes[opened].r);
        }
        uint32_t me = uint32_t(out.size() / kBVHNodeWords);
        out.resize(out.size() + kBVHNodeWords, 0.f);
        s.bvhSah += nd.box.area() / rootArea;
        uint32_t refs[4] = {kInvalid, kInvalid, kInvalid, kInvalid};
        Box boxes[4];
        uint32_t maxDepth = depth, maxNeed = 0;
        for (size_t i = 0; i < ch.size(); i++) {
            Ret r = collapse(ch[i], depth + 1);
            refs[i] = r.ref;
            boxes[i] = B.nodes[ch[i]].box;
            maxDepth = std::max(maxDepth, r.depth);
            maxNeed = std::max(maxNeed, r.need);
        }
        float clo[4][3] = {}, chi[4][3] = {};
        for (int i = 0; i < 4; i++) {
            if (refs[i] == kInvalid) continue;
            for (int a = 0; a < 3; a++) {
                clo[i][a] = padLo(boxes[i].lo[a]);
                chi[i][a] = padHi(boxes[i].hi[a]);
            }
        }
        uint32_t w[kBVHNodeWords];
        packBVH4Node(w, clo, chi, refs);
        std::memcpy(&out[kBVHNodeWords * size_t(me)], w, sizeof(w));
        return Ret{me, maxDepth, uint32_t(ch.size() - 1) + maxNeed};
    };
    Ret root;
    if (B.nodes[0].l < 0) {
        // A single leaf: wrap it in an internal root with one child.
        out.resize(kBVHNodeWords, 0.f);
        const N2& nd = B.nodes[0];
        uint32_t refs[4] = {makeLeaf(nd.first, nd.count), kInvalid, kInvalid, kInvalid};
        float clo[4][3] = {}, chi[4][3] = {};
        for (int a = 0; a < 3; a++) { clo[0][a] = padLo(nd.box.lo[a]); chi[0][a] = padHi(nd.box.hi[a]); }
        uint32_t w[kBVHNodeWords];
        packBVH4Node(w, clo, chi, refs);
        std::memcpy(out.data(), w, sizeof(w));
        s.bvhLeaves = 1;
        root = Ret{0, 1, 0};
    } else {
        root = collapse(0, 1);
    }
    s.bvhNodes = uint32_t(out.size() / kBVHNodeWords);
    s.bvhDepth = root.depth;
    s.stackNeed = root.need;

    // Nodes by depth (children are allocated after their parent, so one pass in index order
    // assigns every depth).
    std::vector<uint32_t> depth(s.bvhNodes, 0u);
    uint32_t maxD = 0;
    for (uint32_t n = 0; n < s.bvhNodes; n++) {
        maxD = std::max(maxD, depth[n]);
        for (int i = 0; i < 4; i++) {
            uint32_t c;
            std::memcpy(&c, &out[kBVHNodeWords * size_t(n) + 12 + i], 4);
            if (c != kInvalid && !childIsLeaf(c)) depth[c] = depth[n] + 1;
        }
    }
    s.levelOff.assign(maxD + 2, 0u);
    for (uint32_t n = 0; n < s.bvhNodes; n++) s.levelOff[depth[n] + 1]++;
    for (uint32_t d = 0; d <= maxD; d++) s.levelOff[d + 1] += s.levelOff[d];
    s.levelNodes.assign(s.bvhNodes, 0u);
    std::vector<uint32_t> fill(s.levelOff.begin(), s.levelOff.end() - 1);
    for (uint32_t n = 0; n < s.bvhNodes; n++) s.levelNodes[fill[depth[n]]++] = n;
}

// ------------------------------------------------------------------ light collection
// Emissive triangles = triangles whose StandardMaterial has nonzero emission, in ascending
// triangle order (LightCollection.cpp:220-262); world-space normal and area as in
// Scene.slang:392-407; flux = luminance(Le) * area * pi (FinalizeIntegration.cs.slang:108);
// active list = flux > 0 (LightCollection.cpp:494-548).
namespace {
// Area of the triangle (texel-space vertices) inside the unit square [x0, x0 + 1] x [y0, y0 + 1]:
// the triangle clipped against the square's four half-planes, then the shoelace formula.
double clippedArea(const double v[3][2], double x0, double y0) {
    double poly[16][2], tmp[16][2];
    int n = 3;
    for (int i = 0; i < 3; i++) { poly[i][0] = v[i][0]; poly[i][1] = v[i][1]; }
    for (int side = 0; side < 4 && n > 0; side++) {
        const int axis = side >> 1;
        const double lim = (side & 1) ? (axis ? y0 + 1.0 : x0 + 1.0) : (axis ? y0 : x0);
        auto inside = [&](const double* p) { return (side & 1) ? p[axis] <= lim : p[axis] >= lim; };
        int m = 0;
        for (int i = 0; i < n; i++) {
            const double* a = poly[i];
            const double* b = poly[(i + 1) % n];
            const bool ia = inside(a), ib = inside(b);
            if (ia) { tmp[m][0] = a[0]; tmp[m][1] = a[1]; m++; }
            if (ia != ib) {
                const double t = (lim - a[axis]) / (b[axis] - a[axis]);
                tmp[m][0] = a[0] + t * (b[0] - a[0]);
                tmp[m][1] = a[1] + t * (b[1] - a[1]);
                m++;
            }
        }
        n = m;
        for (int i = 0; i < n; i++) { poly[i][0] = tmp[i][0]; poly[i][1] = tmp[i][1]; }
    }
    double a2 = 0.0;
    for (int i = 0; i < n; i++) a2 += poly[i][0] * poly[(i + 1) % n][1] - poly[(i + 1) % n][0] * poly[i][1];
    return std::fabs(0.5 * a2);
}

// Average emissive texel value over a triangle's texture footprint (the emissive integrator of
// EmissiveIntegrator.3d.slang and FinalizeIntegration.cs.slang): texel-space vertices
// (uv - floor(uvMin)) * (W, H); every texel the triangle overlaps adds its nearest-sampled value
// (wrap addressing) weighted by the covered area (at most 1); with no coverage (a triangle
// degenerate in texture space) the mean of the texels at the three vertices. Summed in double
// where the reference sums 29.35 fixed point.
f3 averageEmission(const HostScene& s, int tex, const float uv[3][2]) {
    const int32_t* td = &s.texDesc[4 * size_t(tex)];
    const int W = td[1], H = td[2];
    const float* T = &s.texels[size_t(td[0])];
    auto texel = [&](long i, long j) {
        const long x = ((i % W) + W) % W, y = ((j % H) + H) % H;
        const float* q = T + 3 * (size_t(y) * W + x);
        return mk3(q[0], q[1], q[2]);
    };
    const float umin = std::min(uv[0][0], std::min(uv[1][0], uv[2][0]));
    const float vmin = std::min(uv[0][1], std::min(uv[1][1], uv[2][1]));
    const float ou = std::floor(umin), ov = std::floor(vmin);
    double v[3][2];
    double lo[2] = {1e300, 1e300}, hi[2] = {-1e300, -1e300};
    for (int k = 0; k < 3; k++) {
        v[k][0] = double((uv[k][0] - ou) * float(W));
        v[k][1] = double((uv[k][1] - ov) * float(H));
        for (int a = 0; a < 2; a++) { lo[a] = std::min(lo[a], v[k][a]); hi[a] = std::max(hi[a], v[k][a]); }
    }
    double sum[3] = {0, 0, 0}, wsum = 0.0;
    for (long j = long(std::floor(lo[1])); j < long(std::ceil(hi[1])); j++)
        for (long i = long(std::floor(lo[0])); i < long(std::ceil(hi[0])); i++) {
            const double w = std::min(clippedArea(v, double(i), double(j)), 1.0);
            if (!(w > 0.0)) continue;
            const f3 c = texel(i, j);
            sum[0] += w * c.x; sum[1] += w * c.y; sum[2] += w * c.z;
            wsum += w;
        }
    if (wsum > 0.0) return mk3(float(sum[0] / wsum), float(sum[1] / wsum), float(sum[2] / wsum));
    f3 acc = mk3(0.f);
    for (int k = 0; k < 3; k++) {
        const float fu = uv[k][0] - std::floor(uv[k][0]), fv = uv[k][1] - std::floor(uv[k][1]);
        acc = acc + texel(long(fu * float(W)), long(fv * float(H)));
    }
    return acc / 3.f;
}
}  // namespace

// Emission words of the light records (16-19) from the current materials: Le, or for a textured
// emitter emissiveFactor and the texture + 1. Positions, normals, areas and flux are kept.
void refreshLightEmission(HostScene& s) {
    for (uint32_t i = 0; i < s.nLights; i++) {
        float* rec = &s.lights[20 * size_t(i)];
        uint32_t t;
        std::memcpy(&t, &rec[11], 4);
        const float* m = &s.mats[size_t(s.triMat[t]) * M_STRIDE];
        const int etex = int(m[M_EMIT_TEX]) - 1;
        const f3 Le = etex >= 0 ? mk3(m[M_EMIT_FACTOR]) : mk3(m[M_EMIT], m[M_EMIT + 1], m[M_EMIT + 2]) * m[M_EMIT_FACTOR];
        rec[16] = Le.x; rec[17] = Le.y; rec[18] = Le.z;
        rec[19] = etex >= 0 ? float(etex + 1) : 0.f;
    }
}

void buildLights(HostScene& s) {
    s.lights.clear();
    s.activeLights.clear();
    s.triToLight.assign(s.nT, kInvalid);
    uint32_t n = 0;
    for (uint32_t t = 0; t < s.nT; t++) {
        const float* m = &s.mats[size_t(s.triMat[t]) * M_STRIDE];
        if (int(m[M_TYPE]) != MAT_STANDARD) continue;
        const int etex = int(m[M_EMIT_TEX]) - 1;
        f3 Le = mk3(m[M_EMIT], m[M_EMIT + 1], m[M_EMIT + 2]) * m[M_EMIT_FACTOR];
        // Light record radiance: Le, or for a textured emitter emissiveFactor (the texture is
        // evaluated where the light is sampled) with the texture in the fourth word.
        f3 recLe = Le;
        float recTex = 0.f;
        if (etex >= 0) {
            if (!(m[M_EMIT_FACTOR] > 0.f)) continue;
            float uvs[3][2];
            for (int j = 0; j < 3; j++) {
                const uint32_t vi = s.idx[3 * size_t(t) + j];
                uvs[j][0] = s.uv[2 * size_t(vi)];
                uvs[j][1] = s.uv[2 * size_t(vi) + 1];
            }
            Le = averageEmission(s, etex, uvs) * m[M_EMIT_FACTOR];
            recLe = mk3(m[M_EMIT_FACTOR]);
            recTex = float(etex + 1);
        } else if (!anyPos(Le)) {
            continue;
        }
        f3 p[3];
        for (int j = 0; j < 3; j++) {
            const float* q = &s.pos[3 * size_t(s.idx[3 * size_t(t) + j])];
            p[j] = mk3(q[0], q[1], q[2]);
        }
        f3 N = cross(p[1] - p[0], p[2] - p[0]);
        float area = 0.5f * length(N);
        f3 nn = normalize(N);
        float flux = luminance(Le) * area * kPi;
        uint32_t tb = t;
        float tbf;
        std::memcpy(&tbf, &tb, 4);
        float rec[20] = {p[0].x, p[0].y, p[0].z, area, p[1].x, p[1].y, p[1].z, flux, p[2].x, p[2].y, p[2].z, tbf,
                         nn.x, nn.y, nn.z, 0.f, recLe.x, recLe.y, recLe.z, recTex};
        s.lights.insert(s.lights.end(), rec, rec + 20);
        s.triToLight[t] = n;
        if (flux > 0.f) s.activeLights.push_back(n);
        n++;
    }
    s.nLights = n;
}

// ------------------------------------------------------------------ light BVH
// Binned SAOH light BVH (LightBVHBuilder.cpp:228-882; options LightBVHBuilder.h:68-97:
// 16 bins, <= 10 lights per leaf, leaves as soon as possible, pre-integrated flux, lighting
// cones). Nodes are stored depth-first with the left child right after its parent.
namespace {
constexpr float kBadCos = -1.f;  // LightBVHTypes.slang:43
struct LBox {
    f3 lo = mk3(1e30f), hi = mk3(-1e30f);
    bool valid() const { return lo.x <= hi.x && lo.y <= hi.y && lo.z <= hi.z; }
    void grow(const LBox& o) {
        lo = mk3(std::min(lo.x, o.lo.x), std::min(lo.y, o.lo.y), std::min(lo.z, o.lo.z));
        hi = mk3(std::max(hi.x, o.hi.x), std::max(hi.y, o.hi.y), std::max(hi.z, o.hi.z));
    }
    f3 center() const { return (lo + hi) * 0.5f; }
    f3 extent() const { return hi - lo; }
    float area() const { f3 e = extent(); return (e.x * e.y + e.x * e.z + e.y * e.z) * 2.f; }
};
float clampAcos(float v) { return std::acos(std::min(std::max(v, -1.0f), 1.0f)); }
float sinOfCos(float c) { return std::sqrt(std::max(0.f, 1.f - c * c)); }
// Grow a cone (dir, cos) so it also holds another cone (LightBVHBuilder.cpp:67-88).
float growCone(f3 dir, float cosT, f3 odir, float ocos) {
    if (cosT == kBadCos || ocos == kBadCos) return kBadCos;
    float cd = dot(dir, odir);
    float sd = sinOfCos(cd), so = sinOfCos(ocos);
    float ct = ocos * cd - so * sd;
    float st = so * cd + ocos * sd;
    return st > 0.f ? std::min(cosT, ct) : kBadCos;
}
// LightBVHBuilder.cpp:94-109
f3 mergeCones(f3 a, float ac, f3 b, float bc, float& outCos) {
    f3 d = a + b;
    if (ac == kBadCos || bc == kBadCos || (d.x == 0.f && d.y == 0.f && d.z == 0.f)) { outCos = kBadCos; return mk3(0.f); }
    d = normalize(d);
    const float da = clampAcos(dot(d, a)), db = clampAcos(dot(d, b));
    outCos = std::cos(std::max(da + std::acos(ac), db + std::acos(bc)));
    return d;
}
// Conty & Kulla 2018 Eq. 1 for flat emitters (LightBVHBuilder.cpp:659-665).
float orientCost(float to) {
    float tw = std::min(to + kPi2, kPi);
    float st = std::sin(to), ct = std::cos(to);
    return kTwoPi * (1.0f - ct) + kPi2 * (2.0f * tw * st - std::cos(to - 2.0f * tw) - 2.0f * to * st + ct);
}
float saoh(const LBox& b, float flux, float cosT) {
    float a = b.valid() ? b.area() : 0.f;
    float th = cosT != kBadCos ? clampAcos(cosT) : kPi;
    return flux * a * orientCost(th);
}
struct LItem { LBox box; f3 dir; float flux; uint32_t light; float c[3]; };
struct LBuild {
    HostScene& s;
    std::vector<LItem> items;
    std::vector<float>& nodes;
    static constexpr uint32_t kLeaf = 10, kBins = 16;
    float nodeFlux = 0.f;

    f3 coneOf(uint32_t b, uint32_t e, float& cosT) {
        f3 sum = mk3(0.f), dir = mk3(0.f);
        cosT = kBadCos;
        for (uint32_t i = b; i < e; i++) sum = sum + items[i].dir;
        if (length(sum) >= kFltMin) {
            dir = normalize(sum);
            cosT = 1.f;
            for (uint32_t i = b; i < e; i++) cosT = growCone(dir, cosT, items[i].dir, 1.f);
        }
        return dir;
    }
    struct Bin { LBox box; uint32_t n = 0; float flux = 0.f; f3 dir = mk3(0.f); float cosT = 1.f; };
    float sideCost(const Bin* bins, uint32_t from, uint32_t to, Bin& total) {
        float cosT = kBadCos;
        if (length(total.dir) >= kFltMin) {
            cosT = 1.f;
            f3 cd = normalize(total.dir);
            for (uint32_t j = from; j <= to; j++) cosT = growCone(cd, cosT, bins[j].dir, bins[j].cosT);
        }
        return saoh(total.box, total.flux, cosT);
    }
    bool chooseSplit(uint32_t b, uint32_t e, const LBox& nb, uint32_t& axis, uint32_t& at) {
        float best = INFINITY;
        bool found = false;
        f3 dims = nb.extent();
        uint32_t big = dims.z >= dims.x && dims.z >= dims.y ? 2 : (dims.y >= dims.x && dims.y >= dims.z ? 1 : 0);
        for (uint32_t d = 0; d < 3; d++) {
            Bin bins[kBins];
            float lo = comp(nb.lo, int(d)), w = comp(nb.hi, int(d)) - lo;
            float scale = w > kFltMin ? float(kBins) / w : 0.f;
            auto binOf = [&](const LItem& it) { return std::min(uint32_t((it.c[d] - lo) * scale), kBins - 1); };
            for (uint32_t i = b; i < e; i++) {
                Bin& x = bins[binOf(items[i])];
                x.box.grow(items[i].box); x.n += 1; x.flux += items[i].flux; x.dir = x.dir + items[i].dir;
            }
            for (Bin& x : bins) {
                x.cosT = length(x.dir) < kFltMin ? kBadCos : 1.0f;
                x.dir = normalize(x.dir);
            }
            for (uint32_t i = b; i < e; i++) {
                Bin& x = bins[binOf(items[i])];
                x.cosT = growCone(x.dir, x.cosT, items[i].dir, 1.f);
            }
            float cost[kBins - 1];
            Bin left;
            for (uint32_t i = 0; i < kBins - 1; i++) {
                left.box.grow(bins[i].box); left.n += bins[i].n; left.flux += bins[i].flux; left.dir = left.dir + bins[i].dir;
                cost[i] = sideCost(bins, 0, i, left);
            }
            Bin right;
            for (uint32_t i = kBins - 1; i > 0; i--) {
                right.box.grow(bins[i].box); right.n += bins[i].n; right.flux += bins[i].flux; right.dir = right.dir + bins[i].dir;
                cost[i - 1] += sideCost(bins, i, kBins - 1, right);
            }
            float axisBest = INFINITY;
            uint32_t axisAt = b;
            for (uint32_t i = 0, cnt = b; i < kBins - 1; i++) {
                cnt += bins[i].n;
                if (cost[i] < axisBest) { axisBest = cost[i]; axisAt = cnt; }
            }
            axisBest *= comp(dims, int(big)) / comp(dims, int(d));
            if (axisAt == b || axisAt == e) continue;
            if (axisBest < best) { best = axisBest; axis = d; at = axisAt; found = true; }
        }
        if (!found) {
            if (e - b <= kLeaf) return false;
            axis = dims.z >= dims.x && dims.z >= dims.y ? 2 : (dims.y >= dims.x ? 1 : 0);
            at = (b + e) / 2;
        }
        return true;
    }
    uint32_t emit(uint64_t mask, uint32_t depth, uint32_t b, uint32_t e) {
        float flux = 0.f;
        LBox nb;
        for (uint32_t i = b; i < e; i++) { nb.grow(items[i].box); flux += items[i].flux; }
        nodeFlux = flux;
        uint32_t axis = 0, at = 0;
        bool split = (e - b) > kLeaf && chooseSplit(b, e, nb, axis, at);
        uint32_t me = uint32_t(nodes.size() / 16);
        nodes.resize(nodes.size() + 16, 0.f);
        f3 o = (nb.hi + nb.lo) * 0.5f, x = (nb.hi - nb.lo) * 0.5f;
        float* q = &nodes[16 * size_t(me)];
        q[0] = o.x; q[1] = o.y; q[2] = o.z; q[3] = flux;
        q[4] = x.x; q[5] = x.y; q[6] = x.z;
        if (split) {
            std::sort(items.begin() + b, items.begin() + e, [axis](const LItem& p, const LItem& r) {
                return p.c[axis] < r.c[axis] || (p.c[axis] == r.c[axis] && p.light < r.light);
            });
            if (depth >= 64) throw std::runtime_error("light BVH deeper than 64 levels");
            emit(mask, depth + 1, b, at);
            uint32_t right = emit(mask | (1ull << depth), depth + 1, at, e);
            q = &nodes[16 * size_t(me)];
            std::memcpy(&q[12], &right, 4);
        } else {
            float cosT;
            f3 d = coneOf(b, e, cosT);
            q[7] = cosT; q[8] = d.x; q[9] = d.y; q[10] = d.z;
            uint32_t info[4] = {kLeafBit, uint32_t(s.lbvhTri.size()), e - b, 0u};
            std::memcpy(&q[12], info, 16);
            for (uint32_t i = b; i < e; i++) {
                s.lbvhTri.push_back(items[i].light);
                s.lbvhMask[2 * items[i].light] = uint32_t(mask);
                s.lbvhMask[2 * items[i].light + 1] = uint32_t(mask >> 32);
            }
        }
        return me;
    }
    // Bottom-up cone union (LightBVHBuilder.cpp:440-473).
    f3 cones(uint32_t ni, float& cosT) {
        float* q = &nodes[16 * size_t(ni)];
        uint32_t info;
        std::memcpy(&info, &q[12], 4);
        if (!(info & kLeafBit)) {
            float lc, rc;
            f3 ld = cones(ni + 1, lc);
            f3 rd = cones(info, rc);
            f3 d = mergeCones(ld, lc, rd, rc, cosT);
            q = &nodes[16 * size_t(ni)];
            q[7] = cosT; q[8] = d.x; q[9] = d.y; q[10] = d.z;
            return d;
        }
        cosT = q[7];
        return mk3(q[8], q[9], q[10]);
    }
};
}  // namespace

void buildLightBVH(HostScene& s) {
    s.lbvh.clear();
    s.lbvhTri.clear();
    s.lbvhMask.assign(2 * size_t(s.nLights), 0xffffffffu);
    LBuild L{s, {}, s.lbvh};
    for (uint32_t i = 0; i < s.nLights; i++) {
        const float* r = &s.lights[20 * size_t(i)];
        if (!(r[7] > 0.f)) continue;
        LItem it;
        for (int j = 0; j < 3; j++) {
            LBox pb;
            pb.lo = pb.hi = mk3(r[4 * j], r[4 * j + 1], r[4 * j + 2]);
            it.box.grow(pb);
        }
        f3 c = it.box.center();
        it.c[0] = c.x; it.c[1] = c.y; it.c[2] = c.z;
        it.dir = mk3(r[12], r[13], r[14]);
        it.flux = r[7];
        it.light = i;
        L.items.push_back(it);
    }
    if (!L.items.empty()) {
        L.emit(0ull, 0, 0, uint32_t(L.items.size()));
        float c;
        L.cones(0, c);
    }
    s.nLbvh = uint32_t(s.lbvh.size() / 16);
    // Pack the nodes (LightBVHTypes.slang:97-198), then their sampling records after all packed
    // nodes; the host mirror keeps the decoded attributes, which are what sampling sees.
    s.lbvhPacked.assign((8 + kLNodeSampleWords) * size_t(s.nLbvh), 0u);
    for (uint32_t n = 0; n < s.nLbvh; n++) {
        float* q = &s.lbvh[16 * size_t(n)];
        uint32_t info[3];
        std::memcpy(info, &q[12], 12);
        uint32_t w0 = info[0];
        if (info[0] & kLeafBit) {
            if (info[2] > kLNodeCountMask || info[1] > kLNodeOffsetMask)
                throw std::runtime_error("light BVH leaf beyond the packed node's triangle count / offset range");
            w0 = kLeafBit | (info[2] << kLNodeCountShift) | info[1];
        }
        LNodeAttr a;
        a.origin = mk3(q[0], q[1], q[2]);
        a.flux = q[3];
        a.extent = mk3(q[4], q[5], q[6]);
        a.cosCone = q[7];
        a.coneDir = mk3(q[8], q[9], q[10]);
        uint32_t* w = &s.lbvhPacked[8 * size_t(n)];
        packLightNode(w, w0, a);
        const LNodeAttr d = unpackLightNode(make_uint4(w[0], w[1], w[2], w[3]), make_uint4(w[4], w[5], w[6], w[7]));
        float rec[kLNodeSampleWords];
        lightNodeSampleRecord(rec, w[0], d);
        std::memcpy(&s.lbvhPacked[8 * size_t(s.nLbvh) + kLNodeSampleWords * size_t(n)], rec, sizeof(rec));
        q[4] = d.extent.x; q[5] = d.extent.y; q[6] = d.extent.z;
        q[7] = d.cosCone;
        q[8] = d.coneDir.x; q[9] = d.coneDir.y; q[10] = d.coneDir.z;
    }
    // Refit schedule: depth of every node (nodes are depth-first, left child = parent + 1).
    s.lrefitNodes.clear();
    s.lrefitOff.assign(2, 0u);  // no internal level, an empty leaf group
    s.lrefitDepth = 0;
    if (s.nLbvh == 0) return;
    std::vector<uint32_t> depth(s.nLbvh, 0u);
    std::vector<uint8_t> leaf(s.nLbvh, 0);
    for (uint32_t n = 0; n < s.nLbvh; n++) {
        uint32_t info;
        std::memcpy(&info, &s.lbvh[16 * size_t(n) + 12], 4);
        if (info & kLeafBit) { leaf[n] = 1; continue; }
        depth[n + 1] = depth[n] + 1;
        depth[info] = depth[n] + 1;
        s.lrefitDepth = std::max(s.lrefitDepth, depth[n] + 1);
    }
    s.lrefitOff.assign(s.lrefitDepth + 2, 0u);
    for (uint32_t n = 0; n < s.nLbvh; n++) s.lrefitOff[(leaf[n] ? s.lrefitDepth : depth[n]) + 1]++;
    for (uint32_t d = 0; d <= s.lrefitDepth; d++) s.lrefitOff[d + 1] += s.lrefitOff[d];
    s.lrefitNodes.assign(s.nLbvh, 0u);
    std::vector<uint32_t> fill(s.lrefitOff.begin(), s.lrefitOff.end() - 1);
    for (uint32_t n = 0; n < s.nLbvh; n++) s.lrefitNodes[fill[leaf[n] ? s.lrefitDepth : depth[n]]++] = n;
}

void decodeLightNodes(HostScene& s) {
    for (uint32_t n = 0; n < s.nLbvh; n++) {
        const uint32_t* w = &s.lbvhPacked[8 * size_t(n)];
        const LNodeAttr d = unpackLightNode(make_uint4(w[0], w[1], w[2], w[3]), make_uint4(w[4], w[5], w[6], w[7]));
        float* q = &s.lbvh[16 * size_t(n)];
        q[0] = d.origin.x; q[1] = d.origin.y; q[2] = d.origin.z; q[3] = d.flux;
        q[4] = d.extent.x; q[5] = d.extent.y; q[6] = d.extent.z; q[7] = d.cosCone;
        q[8] = d.coneDir.x; q[9] = d.coneDir.y; q[10] = d.coneDir.z; q[11] = 0.f;
        uint32_t info[4] = {w[0], 0u, 0u, 0u};
        if (w[0] & kLeafBit) info[0] = kLeafBit, info[1] = w[0] & kLNodeOffsetMask, info[2] = (w[0] >> kLNodeCountShift) & kLNodeCountMask;
        std::memcpy(&q[12], info, 16);
    }
}

namespace {
int sceneFail(std::string& err, int code, const std::string& msg) {
    err = msg;
    return code;
}
}  // namespace

// fhp_scene_create's validation and host copy of a scene description (FHP_OK, or FHP_EINVAL /
// FHP_EBUILD with the message in err).
int sceneFromDesc(const FhpSceneDesc& d, HostScene& h, std::string& err) {
    if (!d.h_positions || !d.h_normals || !d.h_texcoords || !d.h_indices || !d.h_tri_material || !d.h_materials ||
        !d.h_camera)
        return sceneFail(err, FHP_EINVAL, "scene description has a null array");
    if (d.num_triangles == 0 || d.num_vertices == 0) return sceneFail(err, FHP_EBUILD, "scene has no triangles");
    if (d.num_textures && (!d.h_tex_desc || !d.h_texels)) return sceneFail(err, FHP_EINVAL, "texture arrays missing");
    h.nV = d.num_vertices; h.nT = d.num_triangles; h.nM = d.num_materials; h.nTex = d.num_textures;
    h.pos.assign(d.h_positions, d.h_positions + 3 * size_t(h.nV));
    h.nrm.assign(d.h_normals, d.h_normals + 3 * size_t(h.nV));
    h.restPos = h.pos;
    h.restNrm = h.nrm;
    h.uv.assign(d.h_texcoords, d.h_texcoords + 2 * size_t(h.nV));
    h.idx.assign(d.h_indices, d.h_indices + 3 * size_t(h.nT));
    h.triMat.assign(d.h_tri_material, d.h_tri_material + h.nT);
    h.mats.assign(d.h_materials, d.h_materials + size_t(h.nM) * M_STRIDE);
    for (uint32_t i : h.idx)
        if (i >= h.nV) return sceneFail(err, FHP_EINVAL, "vertex index out of range");
    for (uint32_t m : h.triMat)
        if (m >= h.nM) return sceneFail(err, FHP_EINVAL, "material index out of range");
    if (h.nTex) {
        h.texDesc.assign(d.h_tex_desc, d.h_tex_desc + 4 * size_t(h.nTex));
        h.texels.assign(d.h_texels, d.h_texels + d.num_texels_floats);
        for (uint32_t t = 0; t < h.nTex; t++) {
            int64_t off = h.texDesc[4 * t], w = h.texDesc[4 * t + 1], hh = h.texDesc[4 * t + 2];
            if (off < 0 || w <= 0 || hh <= 0 || uint64_t(off + 3 * w * hh) > d.num_texels_floats)
                return sceneFail(err, FHP_EINVAL, "texture descriptor out of range");
        }
    }
    for (uint32_t m = 0; m < h.nM; m++) {
        int tex = int(h.mats[size_t(m) * M_STRIDE + M_TEX]);
        if (tex >= int(h.nTex)) return sceneFail(err, FHP_EINVAL, "material texture index out of range");
        if (int(h.mats[size_t(m) * M_STRIDE + M_EMIT_TEX]) - 1 >= int(h.nTex))
            return sceneFail(err, FHP_EINVAL, "material emissive texture index out of range");
    }
    std::memcpy(h.cam, d.h_camera, sizeof(h.cam));
    if (d.num_meshes) {
        if (!d.h_mesh_vertex_offsets) return sceneFail(err, FHP_EINVAL, "mesh vertex offsets missing");
        h.meshOff.assign(d.h_mesh_vertex_offsets, d.h_mesh_vertex_offsets + d.num_meshes + 1);
        if (h.meshOff.front() != 0 || h.meshOff.back() != h.nV) return sceneFail(err, FHP_EINVAL, "mesh vertex offsets must span [0, num_vertices]");
        for (uint32_t m = 0; m < d.num_meshes; m++)
            if (h.meshOff[m] > h.meshOff[m + 1]) return sceneFail(err, FHP_EINVAL, "mesh vertex offsets must not decrease");
    } else {
        h.meshOff = {0u, h.nV};
    }
    if (d.flags & ~uint32_t(FHP_SCENE_USE_COMPRESSED_HIT_INFO)) return sceneFail(err, FHP_EINVAL, "unknown scene flags");
    try {
        computeCamera(h);
        computeHitInfo(h, (d.flags & FHP_SCENE_USE_COMPRESSED_HIT_INFO) != 0);
    } catch (const std::exception& e) {
        return sceneFail(err, FHP_EBUILD, e.what());
    }
    return FHP_OK;
}


bool hasTransmission(const HostScene& h) {
    for (uint32_t m = 0; m < h.nM; m++) {
        const float* r = &h.mats[size_t(m) * M_STRIDE];
        if (int(r[M_TYPE]) == MAT_STANDARD && (r[M_SPEC_TRANS] > 0.f || r[M_DIFF_TRANS] > 0.f)) return true;
    }
    return false;
}

std::vector<float> packVertexRecords(const HostScene& h) {
    std::vector<float> vtx(8 * size_t(h.nV));
    for (uint32_t i = 0; i < h.nV; i++) {
        float* q = &vtx[8 * size_t(i)];
        q[0] = h.pos[3 * i]; q[1] = h.pos[3 * i + 1]; q[2] = h.pos[3 * i + 2]; q[3] = h.nrm[3 * i];
        q[4] = h.nrm[3 * i + 1]; q[5] = h.nrm[3 * i + 2]; q[6] = h.uv[2 * i]; q[7] = h.uv[2 * i + 1];
    }
    return vtx;
}

}  // namespace fhp
