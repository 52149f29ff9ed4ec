// Device-resident scene layout and shading routines of the fhp core.
//
// HBM layout (all arrays 16-byte aligned, built once per scene by fhp_scene_create):
//   bvh      BVH4 nodes, 64 B each = 4 x float4, child boxes quantised to 8-bit plane codes
//            (layout at packBVH4Node).
//   triPos   48 B per triangle in BVH leaf order: (v0, origTri) (v1, 0) (v2, 0).
//   idx      3 x u32 per triangle (original order); vtx 32 B per vertex: (p.xyz, n.x) (n.yz, uv).
//   triMat   u32 material id per triangle; mats 80 B per material (20 floats, see MatField).
//   texels   RGB fp32 texels of all textures; texDesc int4 (offset, width, height, 0).
//   lights   80 B per emissive triangle; lbvh 32 B per light-BVH node (packed), then 48 B per
//            node of sampling records (lightNodeSampleRecord).
#pragma once
#include "fhp_math.h"

namespace fhp {

// Material record (20 floats), identical to the scene description the Python layer builds.
enum MatField { M_TYPE = 0, M_BASE = 1, M_ROUGH = 4, M_METAL = 5, M_IOR = 6, M_EMIT = 7, M_EMIT_FACTOR = 10,
                M_DOUBLE_SIDED = 11, M_TEX = 12, M_EMIT_TEX = 13, M_SPEC_TRANS = 14, M_DIFF_TRANS = 15, M_TRANS = 16,
                M_THIN = 19, M_STRIDE = 20 };
// M_EMIT_TEX holds the emissive texture index + 1 (0: the emissive colour is used). M_SPEC_TRANS,
// M_DIFF_TRANS: specular / diffuse transmission; M_TRANS: transmission colour (rgb); M_THIN: thin
// surface (the transmission albedo is the colour itself rather than its square root).
enum MatType { MAT_STANDARD = 0, MAT_PBRT_DIFFUSE = 1 };

// BVH4 child reference: bit 31 = leaf, bits 28..30 = count - 1, bits 0..27 = first triangle.
constexpr uint32_t kLeafBit = 0x80000000u;
FHP_HD bool childIsLeaf(uint32_t c) { return (c & kLeafBit) != 0u; }
FHP_HD uint32_t leafFirst(uint32_t c) { return c & 0x0fffffffu; }
FHP_HD uint32_t leafCount(uint32_t c) { return ((c >> 28) & 7u) + 1u; }
FHP_HD uint32_t makeLeaf(uint32_t first, uint32_t count) { return kLeafBit | ((count - 1u) << 28) | first; }
// BVH4 node, 64 bytes (16 words): words 0-2 the origin o of the node's box, words 3-5 the
// quantisation step s per axis, words 6-11 the children's boxes as 8-bit multiples of s from o
// (word 6 + 2a the four children's lower planes on axis a, byte i for child i; word 7 + 2a the
// upper planes), words 12-15 the child references. A child's decoded plane is fmaf(float(q), s, o)
// in fp32 (the same on host and device); packBVH4Node chooses every code with that arithmetic, so a
// decoded box always contains the child's box (lower planes rounded down, upper planes up).
constexpr uint32_t kBVHNodeWords = 16;
FHP_HD float bvhPlane(float o, float s, uint32_t q) { return fmaf(float(q), s, o); }
FHP_HD void packBVH4Node(uint32_t w[16], const float lo[4][3], const float hi[4][3], const uint32_t ref[4]) {
    float nlo[3] = {INFINITY, INFINITY, INFINITY}, nhi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = 0; i < 4; i++) {
        if (ref[i] == 0xffffffffu) continue;
        for (int a = 0; a < 3; a++) {
            nlo[a] = fminf(nlo[a], lo[i][a]);
            nhi[a] = fmaxf(nhi[a], hi[i][a]);
        }
    }
    for (int k = 0; k < 16; k++) w[k] = 0u;
    for (int a = 0; a < 3; a++) {
        if (!(nlo[a] <= nhi[a])) nlo[a] = nhi[a] = 0.f;  // no child
        const float o = nlo[a];
        float st = (nhi[a] - o) / 255.f;
        // The step must reach the top of the box with the decode's own arithmetic.
        while (bvhPlane(o, st, 255u) < nhi[a]) st = st > 0.f ? u2f(f2u(st) + 1u) : 1e-30f;
        w[a] = f2u(o);
        w[3 + a] = f2u(st);
        for (int i = 0; i < 4; i++) {
            if (ref[i] == 0xffffffffu) continue;
            uint32_t ql = 0u, qh = 255u;
            if (st > 0.f) {
                const float fl = floorf((lo[i][a] - o) / st), fh = ceilf((hi[i][a] - o) / st);
                ql = uint32_t(fminf(fmaxf(fl, 0.f), 255.f));
                qh = uint32_t(fminf(fmaxf(fh, 0.f), 255.f));
                while (ql > 0u && bvhPlane(o, st, ql) > lo[i][a]) ql--;
                while (ql < 255u && bvhPlane(o, st, ql + 1u) <= lo[i][a]) ql++;
                while (qh < 255u && bvhPlane(o, st, qh) < hi[i][a]) qh++;
                while (qh > 0u && bvhPlane(o, st, qh - 1u) >= hi[i][a]) qh--;
            }
            w[6 + 2 * a] |= ql << (8 * i);
            w[7 + 2 * a] |= qh << (8 * i);
        }
    }
    for (int i = 0; i < 4; i++) w[12 + i] = ref[i];
}

// What a pop from beyond a traversal stack's capacity returns (the overflow is already reported):
// the leaf of triangle 0 alone, which every scene has. Testing it reads in-bounds memory and can
// only report a true intersection.
constexpr uint32_t kOverflowLeaf = kLeafBit | 0u;

struct CameraDev {
    f3 pos, U, V, W;
    float nearZ, farZ;
};

// PackedHitInfo bit allocation of a scene (HitInfo.cpp:52-106, HitInfo.slang:300-312): the type in
// the top typeBits of word 0, then the instance id above primBits of primitive index when the
// header fits in 32 bits; otherwise word 0 holds type and instance id and word 1 the primitive.
// Compressed (64-bit) records carry the barycentrics as two unorm16 in word 1; uncompressed
// (128-bit) records carry them as floats in words 2 and 3.
struct HitInfoLayout {
    uint32_t compressed;
    uint32_t typeBits, instBits, primBits;
};

struct SceneDev {
    const float4* bvh;           // BVH4 nodes, 4 x float4 each (packBVH4Node)
    const float4* triPos;
    const uint32_t* idx;
    const float4* vtx;
    const uint32_t* triMat;
    const float* mats;
    const float* texels;
    const float* matsS;    // sampling parameters: == mats / texels unless frozen for finite differences
    const float* texelsS;
    const int4* texDesc;
    const float4* lights;        // 5 x float4 per light: (p0, area) (p1, flux) (p2, tri) (normal, 0) (Le, 0)
    const uint32_t* activeLights;
    const uint32_t* triToLight;
    const uint4* lbvh;           // 2 x uint4 per node, the packed layout of packLightNode
    const float4* lbvhS;         // 3 x float4 per node, lightNodeSampleRecord of the packed node
    const uint32_t* lbvhTri;
    const uint2* lbvhMask;
    CameraDev cam;
    uint32_t nTris, nLights, nActive, nLbvh;
    // Meshes as geometry instances of PackedHitInfo: mesh m owns triangles
    // [meshTriOff[m], meshTriOff[m + 1]) (null when some mesh's triangles are not contiguous).
    const uint32_t* meshTriOff;
    uint32_t nMeshes;
    HitInfoLayout hit;
    float4 viewProjZ, viewProjW;  // rows 2 and 3 of the camera's view-projection matrix (depth = z / w)
    // Environment map (EnvMap.slang): lat-long RGB texels (null: none), world-to-local rotation
    // rows, intensity * tint, flags kEnvLight / kEnvBackground.
    const float* envTex;
    uint32_t envW, envH, envFlags;
    f3 envInv0, envInv1, envInv2, envScale;
    // Hierarchical importance map of the env map (EnvMapSampler): square levels of luminance,
    // the r x r level (r = 1, 2, 4, ..., envImpDim) row-major at offset (r * r - 1) / 3; null when
    // there is no env map or its radiance integrates to zero.
    const float* envImp;
    uint32_t envImpDim;
    // Analytic lights (LightData.slang:48-66 subset): 4 float4 per light, (type, intensity rgb)
    // (posW, openingAngle) (dirW, cosOpeningAngle) (penumbraAngle, 0, 0, 0); null when none.
    const float4* alights;
    uint32_t nALights;
    uint32_t hasTransmission;  // 1: some StandardMaterial has specular or diffuse transmission
};
constexpr uint32_t kLightPoint = 0u, kLightDirectional = 1u;  // LightType::Point, ::Directional
constexpr uint32_t kEnvLight = 1u, kEnvBackground = 2u;

// ------------------------------------------------------------------ shading data
// Scene.slang:440-476 and MaterialFactory.slang:78-104 with world-space geometry.
struct Shading {
    f3 posW, V, faceN, N, T, B;  // (T, B, N) is the shading frame
    float uvx, uvy;
    uint32_t mat;
    bool frontFacing;
    FHP_HD f3 toLocal(f3 v) const { return mk3(dot(v, T), dot(v, B), dot(v, N)); }
    FHP_HD f3 fromLocal(f3 v) const { return T * v.x + B * v.y + N * v.z; }
    FHP_HD f3 orientedFaceN() const { return frontFacing ? faceN : -faceN; }
};
FHP_HD void setFrame(Shading& sd, f3 n) {
    // ShadingFrame.slang:76-82/:127-156 without a valid tangent: T = perp_stark(N), B = N x T.
    sd.N = normalize(n);
    sd.T = perpStark(sd.N);
    sd.B = cross(sd.N, sd.T);
}
FHP_HD void loadVertex(const SceneDev& s, uint32_t i, f3& p, f3& n, float& u, float& v) {
    float4 a = s.vtx[2 * i], b = s.vtx[2 * i + 1];
    p = mk3(a.x, a.y, a.z);
    n = mk3(a.w, b.x, b.y);
    u = b.z; v = b.w;
}
// Shading data from a triangle hit with DXR-style barycentrics (b1, b2).
FHP_HD Shading shadingFromHit(const SceneDev& s, uint32_t tri, float b1, float b2, f3 rayDir) {
    Shading sd;
    uint32_t i0 = s.idx[3 * tri], i1 = s.idx[3 * tri + 1], i2 = s.idx[3 * tri + 2];
    f3 p0, p1, p2, n0, n1, n2;
    float u0, v0, u1, v1, u2, v2;
    loadVertex(s, i0, p0, n0, u0, v0);
    loadVertex(s, i1, p1, n1, u1, v1);
    loadVertex(s, i2, p2, n2, u2, v2);
    float b0 = 1.f - b1 - b2;
    f3 p = p0 * b0; p = p + p1 * b1; p = p + p2 * b2;
    f3 n = n0 * b0; n = n + n1 * b1; n = n + n2 * b2;
    float ux = u0 * b0; ux = ux + u1 * b1; ux = ux + u2 * b2;
    float uy = v0 * b0; uy = uy + v1 * b1; uy = uy + v2 * b2;
    sd.faceN = normalize(cross(p1 - p0, p2 - p0));
    sd.posW = p;
    sd.uvx = ux; sd.uvy = uy;
    sd.V = -rayDir;
    setFrame(sd, normalize(n));
    sd.frontFacing = dot(sd.V, sd.faceN) >= 0.f;
    sd.mat = s.triMat[tri];
    return sd;
}
FHP_HD f3 rayOriginFromShading(const Shading& sd, bool viewside = true) {
    // ShadingData.slang:77-80
    return rayOriginOffset(sd.posW, (sd.frontFacing == viewside) ? sd.faceN : -sd.faceN);
}

// Bilinear fetch at mip 0 with wrap addressing (MaterialSystem.slang:308-320, default sampler
// MaterialSystem.cpp:74-77). Texel ids and weights are returned for the adjoint.
struct TexTap {
    uint32_t base;      // float offset of the texture
    uint32_t t[4];      // texel indices (row-major)
    float w[4];
};
FHP_HD float wrap01(float x) { return x - floorf(x); }
FHP_HD int imod(int a, int m) { int r = a % m; return r < 0 ? r + m : r; }
FHP_HD f3 texFetch(const SceneDev& s, const float* texels, int tex, float u, float v, TexTap* tap) {
    int4 td = s.texDesc[tex];
    int W = td.y, H = td.z;
    float x = wrap01(u) * float(W) - 0.5f, y = wrap01(v) * float(H) - 0.5f;
    float fx0 = floorf(x), fy0 = floorf(y);
    float fx = x - fx0, fy = y - fy0;
    int x0 = int(fx0), y0 = int(fy0);
    int xa = imod(x0, W), xb = imod(x0 + 1, W), ya = imod(y0, H), yb = imod(y0 + 1, H);
    const float* base = texels + td.x;
    const float* q00 = base + 3 * (size_t(ya) * W + xa);
    const float* q10 = base + 3 * (size_t(ya) * W + xb);
    const float* q01 = base + 3 * (size_t(yb) * W + xa);
    const float* q11 = base + 3 * (size_t(yb) * W + xb);
    float wx0 = 1.f - fx, wy0 = 1.f - fy;
    f3 top = mk3(q00[0], q00[1], q00[2]) * wx0 + mk3(q10[0], q10[1], q10[2]) * fx;
    f3 bot = mk3(q01[0], q01[1], q01[2]) * wx0 + mk3(q11[0], q11[1], q11[2]) * fx;
    if (tap) {
        tap->base = uint32_t(td.x);
        tap->t[0] = uint32_t(ya * W + xa); tap->t[1] = uint32_t(ya * W + xb);
        tap->t[2] = uint32_t(yb * W + xa); tap->t[3] = uint32_t(yb * W + xb);
        tap->w[0] = wx0 * wy0; tap->w[1] = fx * wy0; tap->w[2] = wx0 * fy; tap->w[3] = fx * fy;
    }
    return top * wy0 + bot * fy;
}

// Bilinear lookup of the lat-long texels at (u, v), mip 0, wrapping in u and clamping in v (the
// env sampler of EnvMap.cpp:131-135), without intensity and tint.
FHP_HD f3 envLookup(const SceneDev& s, float u, float v) {
    const int W = int(s.envW), H = int(s.envH);
    const float x = u * float(W) - 0.5f, y = v * float(H) - 0.5f;
    const float fx0 = floorf(x), fy0 = floorf(y);
    const float fx = x - fx0, fy = y - fy0;
    const int x0 = int(fx0), y0 = int(fy0);
    const int xa = imod(x0, W), xb = imod(x0 + 1, W);
    const int ya = y0 < 0 ? 0 : (y0 >= H ? H - 1 : y0), yb = y0 + 1 < 0 ? 0 : (y0 + 1 >= H ? H - 1 : y0 + 1);
    const float* q00 = s.envTex + 3 * (size_t(ya) * W + xa);
    const float* q10 = s.envTex + 3 * (size_t(ya) * W + xb);
    const float* q01 = s.envTex + 3 * (size_t(yb) * W + xa);
    const float* q11 = s.envTex + 3 * (size_t(yb) * W + xb);
    const float wx0 = 1.f - fx, wy0 = 1.f - fy;
    const f3 top = mk3(q00[0], q00[1], q00[2]) * wx0 + mk3(q10[0], q10[1], q10[2]) * fx;
    const f3 bot = mk3(q01[0], q01[1], q01[2]) * wx0 + mk3(q11[0], q11[1], q11[2]) * fx;
    return top * wy0 + bot * fy;
}

// world_to_latlong_map (MathHelpers.slang:111-118) of a unit direction.
FHP_HD void latlongUV(f3 p, float& u, float& v) {
    u = atan2f(p.x, -p.z) * (0.5f * kInvPi) + 0.5f;
    v = acosf(fminf(fmaxf(p.y, -1.f), 1.f)) * kInvPi;
}

// EnvMap::eval(dir) (EnvMap.slang:58-68): direction to local space, lat-long coordinates,
// bilinear lookup, times intensity * tint.
FHP_HD f3 envEval(const SceneDev& s, f3 dir) {
    const f3 p = normalize(mk3(dot(s.envInv0, dir), dot(s.envInv1, dir), dot(s.envInv2, dir)));
    float u, v;
    latlongUV(p, u, v);
    const f3 c = envLookup(s, u, v);
    return mk3(c.x * s.envScale.x, c.y * s.envScale.y, c.z * s.envScale.z);
}

FHP_HD float signf(float x) { return x > 0.f ? 1.f : (x < 0.f ? -1.f : 0.f); }

// Equal-area octahedral map (Clarberg 2008), oct_to_ndir_equal_area_unorm and
// ndir_to_oct_equal_area_unorm (MathHelpers.slang:211-254): the centre of [0,1]^2 is +z, the
// corners -z.
FHP_HD f3 octToDirEqualArea(float px, float py) {
    px = px * 2.f - 1.f;
    py = py * 2.f - 1.f;
    const float d = 1.f - (fabsf(px) + fabsf(py));
    const float r = 1.f - fabsf(d);
    const float phi = r > 0.f ? ((fabsf(py) - fabsf(px)) / r + 1.f) * (0.25f * kPi) : 0.f;
    const float f = r * sqrtf(2.f - r * r);
    return mk3(f * signf(px) * cosf(phi), f * signf(py) * sinf(phi), signf(d) * (1.f - r * r));
}
FHP_HD void dirToOctEqualArea(f3 n, float& ux, float& uy) {
    const float r = sqrtf(1.f - fabsf(n.z));
    const float phi = atan2f(fabsf(n.y), fabsf(n.x));
    float py = r * phi * (2.f * kInvPi);
    float px = r - py;
    if (n.z < 0.f) {
        const float t = px;
        px = 1.f - py;
        py = 1.f - t;
    }
    px *= signf(n.x);
    py *= signf(n.y);
    ux = px * 0.5f + 0.5f;
    uy = py * 0.5f + 0.5f;
}

FHP_HD f3 envToWorld(const SceneDev& s, f3 l) { return s.envInv0 * l.x + s.envInv1 * l.y + s.envInv2 * l.z; }

// EnvMapSampler::sample (EnvMapSampler.slang:73-145): descend the importance map from the 2x2
// level, choosing a column by its two-texel sums and then a row, rescaling (ux, uy) at each step;
// the final texel plus the remaining (ux, uy) is a point of the octahedral map. pdf is per solid
// angle: texel value / mean / (4 pi).
FHP_HD bool envSample(const SceneDev& s, float ux, float uy, f3& dir, float& pdf, f3& Le) {
    const uint32_t dim = s.envImpDim;
    const float avg = s.envImp[0];
    uint32_t px = 0, py = 0;
    for (uint32_t r = 2; r <= dim; r <<= 1) {
        px *= 2u;
        py *= 2u;
        const float* L = s.envImp + (r * r - 1u) / 3u;
        const float w0 = L[py * r + px], w1 = L[py * r + px + 1u], w2 = L[(py + 1u) * r + px], w3 = L[(py + 1u) * r + px + 1u];
        const float q0 = w0 + w2, q1 = w1 + w3;
        const float d = q0 / (q0 + q1);
        uint32_t ox, oy;
        if (ux < d) { ox = 0u; ux = ux / d; }
        else { ox = 1u; ux = (ux - d) / (1.f - d); }
        const float e = ox == 0u ? (w0 / q0) : (w1 / q1);
        if (uy < e) { oy = 0u; uy = uy / e; }
        else { oy = 1u; uy = (uy - e) / (1.f - e); }
        px += ox;
        py += oy;
    }
    const float inv = 1.f / float(dim);
    const f3 l = octToDirEqualArea((float(px) + ux) * inv, (float(py) + uy) * inv);
    const float* L0 = s.envImp + (dim * dim - 1u) / 3u;
    pdf = L0[py * dim + px] / avg * (0.25f * kInvPi);
    dir = envToWorld(s, l);
    Le = envEval(s, dir);
    return true;
}

// EnvMapSampler::evalPdf (EnvMapSampler.slang:154-163): the finest level at the direction's
// octahedral texel (point sampling, clamped), over the mean, per solid angle.
FHP_HD float envPdf(const SceneDev& s, f3 dir) {
    const f3 l = mk3(dot(s.envInv0, dir), dot(s.envInv1, dir), dot(s.envInv2, dir));
    float ux, uy;
    dirToOctEqualArea(l, ux, uy);
    const int dim = int(s.envImpDim);
    const int x = min(max(int(floorf(ux * float(dim))), 0), dim - 1), y = min(max(int(floorf(uy * float(dim))), 0), dim - 1);
    const float* L0 = s.envImp + (uint32_t(dim) * uint32_t(dim) - 1u) / 3u;
    return L0[y * dim + x] / s.envImp[0] * (0.25f * kInvPi);
}

// ------------------------------------------------------------------ material instance
struct Material {
    int type;
    f3 diffuse, specular, emission, base;
    f3 emitColor;  // emissive colour before emissiveFactor (texture value when textured)
    f3 transAlbedo;                  // transmission albedo (StandardBSDF.slang: sqrt(colour) unless thin)
    float specTrans, diffTrans, eta;  // transmission weights; relative IoR (incident / transmitted side)
    float roughness, metallic, F0;
    bool doubleSided;
    int tex;
};
// StandardMaterial.slang:50-126 / PBRTDiffuseMaterial.slang (reflection lobes only).
FHP_HD Material loadMaterial(const SceneDev& s, const float* mats, const float* texels, const Shading& sd, Shading& frameOut,
                             TexTap* tap) {
    Material m;
    const float* r = mats + size_t(sd.mat) * M_STRIDE;
    m.type = int(r[M_TYPE]);
    m.tex = int(r[M_TEX]);
    f3 base = mk3(r[M_BASE], r[M_BASE + 1], r[M_BASE + 2]);
    if (m.tex >= 0) base = texFetch(s, texels, m.tex, sd.uvx, sd.uvy, tap);
    m.base = base;
    m.doubleSided = r[M_DOUBLE_SIDED] != 0.f;
    m.specTrans = 0.f;
    m.diffTrans = 0.f;
    m.transAlbedo = mk3(0.f);
    m.eta = 1.f;
    if (m.type == MAT_STANDARD) {
        m.specTrans = r[M_SPEC_TRANS];
        m.diffTrans = r[M_DIFF_TRANS];
        // Transmissive materials are double-sided (BasicMaterial::adjustDoubleSidedFlag, :662-673).
        if (m.specTrans > 0.f || m.diffTrans > 0.f) {
            m.doubleSided = true;
            const f3 t = mk3(r[M_TRANS], r[M_TRANS + 1], r[M_TRANS + 2]);
            m.transAlbedo = r[M_THIN] != 0.f ? t : mk3(sqrtf(t.x), sqrtf(t.y), sqrtf(t.z));
        }
        // StandardMaterial.slang:56-57 with the incident medium's IoR 1.
        m.eta = sd.frontFacing ? (1.f / r[M_IOR]) : (r[M_IOR] / 1.f);
    }
    // ShadingUtils.slang:107-117 (flipShadingNormal)
    if (!sd.frontFacing && m.doubleSided) frameOut.N = -frameOut.N;
    if (m.type == MAT_PBRT_DIFFUSE) {
        m.diffuse = base; m.specular = mk3(0.f); m.roughness = 0.f; m.metallic = 0.f; m.F0 = 0.f;
        m.emission = mk3(0.f);
        m.emitColor = mk3(0.f);
        return m;
    }
    float ior = r[M_IOR];
    float f = (ior - 1.f) / (ior + 1.f);
    m.F0 = f * f;
    m.metallic = r[M_METAL];
    m.diffuse = base + (mk3(0.f) - base) * m.metallic;                    // lerp(base, 0, metallic)
    m.specular = mk3(m.F0) + (base - mk3(m.F0)) * m.metallic;             // lerp(F0, base, metallic)
    m.roughness = r[M_ROUGH];
    // MaterialSystem::evalEmissive (MaterialSystem.slang:201-213): the emissive texture (default
    // sampler, mip 0) or colour, times emissiveFactor, on the front side only.
    m.emitColor = mk3(0.f);
    m.emission = mk3(0.f);
    if (sd.frontFacing) {
        const int etex = int(r[M_EMIT_TEX]) - 1;
        m.emitColor = etex >= 0 ? texFetch(s, texels, etex, sd.uvx, sd.uvy, nullptr) : mk3(r[M_EMIT], r[M_EMIT + 1], r[M_EMIT + 2]);
        m.emission = m.emitColor * r[M_EMIT_FACTOR];
    }
    return m;
}

// Fresnel.slang:68-92: dielectric Fresnel reflectance for relative IoR eta; cosT is the cosine
// of the refracted direction (0 on total internal reflection, which returns 1).
FHP_HD float fresnelDielectric(float eta, float cosI, float& cosT) {
    if (cosI < 0.f) {
        eta = 1.f / eta;
        cosI = -cosI;
    }
    const float sin2T = eta * eta * (1.f - cosI * cosI);
    if (sin2T > 1.f) {
        cosT = 0.f;
        return 1.f;
    }
    cosT = sqrtf(1.f - sin2T);
    const float rs = (eta * cosI - cosT) / (eta * cosI + cosT);
    const float rp = (eta * cosT - cosI) / (eta * cosT + cosI);
    return 0.5f * (rs * rs + rp * rp);
}

// StandardBSDF (StandardBSDF.slang:125-313): Frostbite diffuse reflection, Lambert diffuse
// transmission (LambertDiffuseBTDF.slang), GGX specular reflection (SpecularMicrofacetBRDF) and
// GGX dielectric reflection + transmission (SpecularMicrofacetBSDF, SpecularMicrofacet.slang:206-471),
// mixed by specularTransmission / diffuseTransmission and sampled one lobe at a time.
struct Bsdf {
    f3 diff, spec, trans;
    float rough, alpha, alphaT, eta, specTrans, diffTrans;
    float pDiff, pDiffT, pSpec, pSpecT;
    FHP_HD void init(f3 wi, const Material& m) {
        diff = m.diffuse; spec = m.specular; rough = m.roughness;
        trans = m.transAlbedo; eta = m.eta;
        specTrans = m.specTrans; diffTrans = m.diffTrans;
        alpha = rough * rough;
        if (alpha < kMinGGXAlpha) alpha = 0.f;
        alphaT = eta == 1.f ? 0.f : alpha;
        float metallicBRDF = m.metallic * (1.f - specTrans);
        float dielectricBSDF = (1.f - m.metallic) * (1.f - specTrans);
        float specularBSDF = specTrans;
        float dw = luminance(m.diffuse);
        float sw = luminance(schlick3(m.specular, 1.f, wi.z));
        pDiff = dw * dielectricBSDF * (1.f - diffTrans);
        pDiffT = dw * dielectricBSDF * diffTrans;
        pSpec = sw * (metallicBRDF + dielectricBSDF);
        pSpecT = specularBSDF;
        float norm = pDiff + pDiffT + pSpec + pSpecT;
        if (norm > 0.f) {
            norm = 1.f / norm;
            pDiff *= norm;
            pDiffT *= norm;
            pSpec *= norm;
            pSpecT *= norm;
        }
    }
    // FrostbiteDiffuseBRDF.slang:84-96 (returns f * pi without the cosine)
    FHP_HD f3 frostbite(f3 wi, f3 wo) const {
        f3 h = normalize(wi + wo);
        float woDotH = dot(wo, h);
        float energyBias = 0.f + (0.5f - 0.f) * rough;
        float energyFactor = 1.f + (1.f / 1.51f - 1.f) * rough;
        float fd90 = energyBias + 2.f * woDotH * woDotH * rough;
        float wiS = schlick1(1.f, fd90, wi.z);
        float woS = schlick1(1.f, fd90, wo.z);
        return diff * wiS * woS * energyFactor;
    }
    FHP_HD f3 evalDiffuse(f3 wi, f3 wo) const {
        if (fminf_(wi.z, wo.z) < kMinCosTheta) return mk3(0.f);
        return frostbite(wi, wo) * kInvPi * wo.z;
    }
    // LambertDiffuseBTDF.slang:40-84
    FHP_HD f3 evalDiffuseT(f3 wi, f3 wo) const {
        if (fminf_(wi.z, -wo.z) < kMinCosTheta) return mk3(0.f);
        return kInvPi * trans * -wo.z;
    }
    FHP_HD float pdfDiffuseT(f3 wi, f3 wo) const {
        if (fminf_(wi.z, -wo.z) < kMinCosTheta) return 0.f;
        return kInvPi * -wo.z;
    }
    // SpecularMicrofacet.slang:58-83
    FHP_HD f3 evalSpecular(f3 wi, f3 wo) const {
        if (fminf_(wi.z, wo.z) < kMinCosTheta) return mk3(0.f);
        if (alpha == 0.f) return mk3(0.f);
        f3 h = normalize(wi + wo);
        float wiDotH = dot(wi, h);
        float D = ggxD(alpha, h.z);
        float G = ggxG(alpha, wi.z, wo.z);
        f3 F = schlick3(spec, 1.f, wiDotH);
        return F * D * G * 0.25f / wi.z;
    }
    // SpecularMicrofacetBSDF::eval (both the reflection and the transmission lobe active)
    FHP_HD f3 evalSpecularT(f3 wi, f3 wo) const {
        if (fminf_(wi.z, fabsf(wo.z)) < kMinCosTheta) return mk3(0.f);
        if (alphaT == 0.f) return mk3(0.f);
        const bool isR = wo.z > 0.f;
        f3 h = normalize(wo + wi * (isR ? 1.f : eta));
        h = h * (h.z > 0.f ? 1.f : (h.z < 0.f ? -1.f : 0.f));
        const float wiDotH = dot(wi, h), woDotH = dot(wo, h);
        const float D = ggxD(alphaT, h.z);
        const float G = ggxG(alphaT, wi.z, fabsf(wo.z));
        float cosT;
        const float F = fresnelDielectric(eta, wiDotH, cosT);
        if (isR) return mk3(F * D * G * 0.25f / wi.z);
        const float sqrtDenom = woDotH + eta * wiDotH;
        const float t = eta * eta * wiDotH * woDotH / (wi.z * sqrtDenom * sqrtDenom);
        return trans * (1.f - F) * D * G * fabsf(t);
    }
    FHP_HD float pdfDiffuse(f3 wi, f3 wo) const {
        if (fminf_(wi.z, wo.z) < kMinCosTheta) return 0.f;
        return kInvPi * wo.z;
    }
    FHP_HD float pdfSpecular(f3 wi, f3 wo) const {
        if (fminf_(wi.z, wo.z) < kMinCosTheta) return 0.f;
        if (alpha == 0.f) return 0.f;
        f3 h = normalize(wi + wo);
        float wiDotH = dot(wi, h);
        return ggxPdfVndf(alpha, wi, h) / (4.f * wiDotH);
    }
    // SpecularMicrofacetBSDF::evalPdf
    FHP_HD float pdfSpecularT(f3 wi, f3 wo) const {
        if (fminf_(wi.z, fabsf(wo.z)) < kMinCosTheta) return 0.f;
        if (alphaT == 0.f) return 0.f;
        const bool isR = wo.z > 0.f;
        f3 h = normalize(wo + wi * (isR ? 1.f : eta));
        h = h * (h.z > 0.f ? 1.f : (h.z < 0.f ? -1.f : 0.f));
        const float wiDotH = dot(wi, h), woDotH = dot(wo, h);
        float cosT;
        const float F = fresnelDielectric(eta, wiDotH, cosT);
        float pdf = ggxPdfVndf(alphaT, wi, h);
        if (isR) {
            pdf /= 4.f * woDotH;
        } else {
            if (woDotH > 0.f) return 0.f;
            const float sqrtDenom = woDotH + eta * wiDotH;
            const float denom = sqrtDenom * sqrtDenom;
            pdf = (denom > 0.f) ? pdf * fabsf(woDotH) / denom : kFltMax;
        }
        pdf *= isR ? F : 1.f - F;
        return pdf;
    }
    FHP_HD f3 eval(f3 wi, f3 wo) const {
        f3 r = mk3(0.f);
        if (pDiff > 0.f) r = r + evalDiffuse(wi, wo) * ((1.f - specTrans) * (1.f - diffTrans));
        if (pDiffT > 0.f) r = r + evalDiffuseT(wi, wo) * ((1.f - specTrans) * diffTrans);
        if (pSpec > 0.f) r = r + evalSpecular(wi, wo) * (1.f - specTrans);
        if (pSpecT > 0.f) r = r + evalSpecularT(wi, wo) * specTrans;
        return r;
    }
    FHP_HD float evalPdf(f3 wi, f3 wo) const {
        float p = 0.f;
        if (pDiff > 0.f) p += pDiff * pdfDiffuse(wi, wo);
        if (pDiffT > 0.f) p += pDiffT * pdfDiffuseT(wi, wo);
        if (pSpec > 0.f) p += pSpec * pdfSpecular(wi, wo);
        if (pSpecT > 0.f) p += pSpecT * pdfSpecularT(wi, wo);
        return p;
    }
    // SpecularMicrofacet.slang:85-157
    FHP_HD bool sampleSpec(f3 wi, f3& wo, float& pdf, f3& w, bool& delta, Rng& rng) const {
        wo = mk3(0.f); w = mk3(0.f); pdf = 0.f;
        if (wi.z < kMinCosTheta) return false;
        if (alpha == 0.f) {
            wo = mk3(-wi.x, -wi.y, wi.z);
            pdf = 0.f;
            w = schlick3(spec, 1.f, wi.z);
            delta = true;
            return true;
        }
        float ux = rng.u1(), uy = rng.u1();
        f3 h = ggxSampleVndf(alpha, wi, ux, uy, pdf);
        float wiDotH = dot(wi, h);
        wo = 2.f * wiDotH * h - wi;
        if (wo.z < kMinCosTheta) return false;
        float G = ggxG(alpha, wi.z, wo.z);
        float GOverG1wo = G * (1.f + ggxLambda(alpha * alpha, wi.z));
        f3 F = schlick3(spec, 1.f, wiDotH);
        pdf /= (4.f * wiDotH);
        w = F * GOverG1wo;
        return true;
    }
    // SpecularMicrofacetBSDF::sample (SpecularMicrofacet.slang:260-410): reflection or refraction
    // chosen by the Fresnel term, VNDF half vectors, delta lobes when alpha is 0.
    FHP_HD bool sampleSpecT(f3 wi, f3& wo, float& pdf, f3& w, bool& delta, Rng& rng) const {
        wo = mk3(0.f); w = mk3(0.f); pdf = 0.f;
        if (wi.z < kMinCosTheta) return false;
        const float lobeSample = rng.u1();
        if (alphaT == 0.f) {
            float cosT;
            const float F = fresnelDielectric(eta, wi.z, cosT);
            const bool isR = lobeSample < F;
            pdf = 0.f;
            w = isR ? mk3(1.f) : trans;
            wo = isR ? mk3(-wi.x, -wi.y, wi.z) : mk3(-wi.x * eta, -wi.y * eta, -cosT);
            delta = true;
            if (fabsf(wo.z) < kMinCosTheta || ((wo.z > 0.f) != isR)) return false;
            return true;
        }
        const float ux = rng.u1(), uy = rng.u1();
        const f3 h = ggxSampleVndf(alphaT, wi, ux, uy, pdf);
        const float wiDotH = dot(wi, h);
        float cosT;
        const float F = fresnelDielectric(eta, wiDotH, cosT);
        const bool isR = lobeSample < F;
        wo = isR ? (2.f * wiDotH * h - wi) : ((eta * wiDotH - cosT) * h - eta * wi);
        if (fabsf(wo.z) < kMinCosTheta || ((wo.z > 0.f) != isR)) return false;
        const float woDotH = dot(wo, h);
        const float G = ggxG(alphaT, wi.z, fabsf(wo.z));
        const float GOverG1wo = G * (1.f + ggxLambda(alphaT * alphaT, wi.z));
        w = mk3(GOverG1wo);
        if (isR) {
            pdf /= 4.f * woDotH;
        } else {
            const float sqrtDenom = woDotH + eta * wiDotH;
            const float denom = sqrtDenom * sqrtDenom;
            pdf = (denom > 0.f) ? pdf * fabsf(woDotH) / denom : kFltMax;
            w = w * (trans * eta * eta);
        }
        pdf *= isR ? F : 1.f - F;
        return true;
    }
    // StandardBSDF.slang:242-313 (one-sample lobe selection)
    FHP_HD bool sample(f3 wi, f3& wo, float& pdf, f3& w, bool& delta, Rng& rng) const {
        wo = mk3(0.f); w = mk3(0.f); pdf = 0.f; delta = false;
        bool valid = false;
        float uSel = rng.u1();
        if (uSel < pDiff) {
            float ux = rng.u1(), uy = rng.u1();
            wo = cosineHemisphere(ux, uy, pdf);
            if (fminf_(wi.z, wo.z) < kMinCosTheta) { w = mk3(0.f); valid = false; }
            else { w = frostbite(wi, wo); valid = true; }
            w = w / pDiff;
            w = w * ((1.f - specTrans) * (1.f - diffTrans));
            pdf *= pDiff;
            if (pSpec > 0.f) pdf += pSpec * pdfSpecular(wi, wo);
            if (pSpecT > 0.f) pdf += pSpecT * pdfSpecularT(wi, wo);
        } else if (uSel < pDiff + pDiffT) {
            float ux = rng.u1(), uy = rng.u1();
            wo = cosineHemisphere(ux, uy, pdf);
            wo.z = -wo.z;
            if (fminf_(wi.z, -wo.z) < kMinCosTheta) { w = mk3(0.f); valid = false; }
            else { w = trans; valid = true; }
            w = w / pDiffT;
            w = w * ((1.f - specTrans) * diffTrans);
            pdf *= pDiffT;
            if (pSpecT > 0.f) pdf += pSpecT * pdfSpecularT(wi, wo);
        } else if (uSel < pDiff + pDiffT + pSpec) {
            valid = sampleSpec(wi, wo, pdf, w, delta, rng);
            w = w / pSpec;
            w = w * (1.f - specTrans);
            pdf *= pSpec;
            if (pDiff > 0.f) pdf += pDiff * pdfDiffuse(wi, wo);
            if (pSpecT > 0.f) pdf += pSpecT * pdfSpecularT(wi, wo);
        } else if (pSpecT > 0.f) {
            valid = sampleSpecT(wi, wo, pdf, w, delta, rng);
            w = w / pSpecT;
            w = w * specTrans;
            pdf *= pSpecT;
            if (pDiff > 0.f) pdf += pDiff * pdfDiffuse(wi, wo);
            if (pDiffT > 0.f) pdf += pDiffT * pdfDiffuseT(wi, wo);
            if (pSpec > 0.f) pdf += pSpec * pdfSpecular(wi, wo);
        }
        return valid;
    }
};

// ShadingUtils.slang:214-240 (reflection or transmission, fullSphere = false) and :155-181.
FHP_HD bool validRT(const Shading& sd, f3 wiL, f3 woL, f3 wo) {
    bool isT = wiL.z * woL.z < 0.f;
    if (fminf_(wiL.z, fabsf(woL.z)) < kMinCosTheta) return false;
    bool wiTop = sd.frontFacing, woTop = dot(wo, sd.faceN) >= 0.f;
    return isT ? (wiTop != woTop) : (wiTop == woTop);
}
FHP_HD bool validR(const Shading& sd, f3 wiL, f3 woL, f3 wo) {
    if (fminf_(wiL.z, woL.z) < kMinCosTheta) return false;
    return sd.frontFacing == (dot(wo, sd.faceN) >= 0.f);
}
// StandardMaterialInstance.slang:63-74 / PBRTDiffuseMaterialInstance.slang:54-64
FHP_HD f3 matEval(const Shading& sd, const Material& m, f3 wo) {
    f3 wiL = sd.toLocal(sd.V), woL = sd.toLocal(wo);
    if (m.type == MAT_PBRT_DIFFUSE) {
        if (!validR(sd, wiL, woL, wo)) return mk3(0.f);
        return m.diffuse * fmaxf_(woL.z, 0.f) * kInvPi;
    }
    if (!validRT(sd, wiL, woL, wo)) return mk3(0.f);
    Bsdf b;
    b.init(wiL, m);
    return b.eval(wiL, woL);
}
// Evaluation with lobe-selection probabilities taken from another parameter set (the sampling
// parameters of the finite-difference mode); identical to matEval when both sets are equal.
FHP_HD f3 matEvalGated(const Shading& sd, const Material& m, const Material& mg, f3 wo) {
    f3 wiL = sd.toLocal(sd.V), woL = sd.toLocal(wo);
    if (m.type == MAT_PBRT_DIFFUSE) {
        if (!validR(sd, wiL, woL, wo)) return mk3(0.f);
        return m.diffuse * fmaxf_(woL.z, 0.f) * kInvPi;
    }
    if (!validRT(sd, wiL, woL, wo)) return mk3(0.f);
    Bsdf b, g;
    b.init(wiL, m);
    g.init(wiL, mg);
    b.pDiff = g.pDiff;
    b.pDiffT = g.pDiffT;
    b.pSpec = g.pSpec;
    b.pSpecT = g.pSpecT;
    return b.eval(wiL, woL);
}
FHP_HD float matPdf(const Shading& sd, const Material& m, f3 wo) {
    f3 wiL = sd.toLocal(sd.V), woL = sd.toLocal(wo);
    if (m.type == MAT_PBRT_DIFFUSE) return fmaxf_(woL.z, 0.f) * kInvPi;
    if (!validRT(sd, wiL, woL, wo)) return 0.f;
    Bsdf b;
    b.init(wiL, m);
    return b.evalPdf(wiL, woL);
}
struct Scatter { f3 wo, weight; float pdf; bool delta, transmission; };
// StandardMaterialInstance.slang:89-110 (+ sampleReference :202-231) / PBRTDiffuseMaterialInstance.slang:66-78
FHP_HD bool matSample(const Shading& sd, const Material& m, Rng& rng, Scatter& s, bool importance) {
    f3 wiL = sd.toLocal(sd.V), woL;
    s.delta = false;
    s.transmission = false;
    if (m.type == MAT_PBRT_DIFFUSE) {
        float ux = rng.u1(), uy = rng.u1();
        woL = cosineHemisphere(ux, uy, s.pdf);
        s.weight = m.diffuse;
        s.wo = sd.fromLocal(woL);
        return validR(sd, wiL, woL, s.wo) && s.pdf != 0.f;
    }
    Bsdf b;
    if (!importance) {
        // sampleReference (StandardMaterialInstance.slang:202-231): the cosine hemisphere, mirrored
        // below the surface with probability 1/2 when a transmission lobe is active.
        const bool transmissive = m.specTrans > 0.f || (m.diffTrans > 0.f && m.specTrans < 1.f && anyPos(m.diffuse));
        float ux = rng.u1(), uy = rng.u1();
        woL = cosineHemisphere(ux, uy, s.pdf);
        if (transmissive) {
            if (rng.u1() < 0.5f) woL.z = -woL.z;
            s.pdf *= 0.5f;
            if (fminf_(fabsf(wiL.z), fabsf(woL.z)) < kMinCosTheta || s.pdf == 0.f) return false;
        } else if (fminf_(wiL.z, woL.z) < kMinCosTheta || s.pdf == 0.f) {
            return false;
        }
        b.init(wiL, m);
        s.weight = b.eval(wiL, woL) / s.pdf;
    } else {
        b.init(wiL, m);
        if (!b.sample(wiL, woL, s.pdf, s.weight, s.delta, rng)) return false;
    }
    s.transmission = woL.z < 0.f;
    s.wo = sd.fromLocal(woL);
    return validRT(sd, wiL, woL, s.wo);
}

// ------------------------------------------------------------------ emissive lights
struct LightSample { f3 posW, normalW, Le, dir; float pdf, distance; uint32_t light; };
// EmissiveLightSamplerHelpers.slang:54-101
FHP_HD bool sampleTri(const SceneDev& s, f3 posW, uint32_t li, float ux, float uy, LightSample& ls) {
    const float4* L = s.lights + 5 * size_t(li);
    float4 a = L[0], b = L[1], c = L[2], n = L[3], e = L[4];
    ls.light = li;
    f3 bc = triangleSample(ux, uy);
    ls.posW = mk3(a.x, a.y, a.z) * bc.x + mk3(b.x, b.y, b.z) * bc.y + mk3(c.x, c.y, c.z) * bc.z;
    f3 toL = ls.posW - posW;
    const float d2 = fmaxf_(kFltMin, dot(toL, toL));
    ls.distance = sqrtf(d2);
    ls.dir = toL / ls.distance;
    ls.normalW = mk3(n.x, n.y, n.z);
    float cosT = dot(ls.normalW, -ls.dir);
    if (cosT <= 0.f) return false;
    // MaterialSystem::evalEmissive at the sample (EmissiveLightSamplerHelpers.slang:79): the
    // record carries the material's current emission (refreshed on material updates; the flux
    // used for selection stays as built, as in the reference): Le, or for a textured emitter
    // emissiveFactor with the texture evaluated here at the sample's uv.
    ls.Le = mk3(e.x, e.y, e.z);
    if (e.w > 0.f) {
        const uint32_t* tv = s.idx + 3 * size_t(f2u(c.w));
        const float4 t0 = s.vtx[2 * size_t(tv[0]) + 1], t1 = s.vtx[2 * size_t(tv[1]) + 1], t2 = s.vtx[2 * size_t(tv[2]) + 1];
        const float u = t0.z * bc.x + t1.z * bc.y + t2.z * bc.z, v = t0.w * bc.x + t1.w * bc.y + t2.w * bc.z;
        ls.Le = texFetch(s, s.texels, int(e.w) - 1, u, v, nullptr) * ls.Le;
    }
    float denom = fmaxf_(kFltMin, cosT * a.w);
    ls.pdf = d2 / denom;
    return true;
}
// MinimalPathTracer.rt.slang:253-286 (evalDirectAnalytic) with LightHelpers.slang:190-258: one
// analytic light picked uniformly, its unshadowed contribution mi.eval * Li * lightCount, and the
// shadow ray that gates it. Returns false when the sample is rejected before the shadow ray.
FHP_HD float smoothstep01(float e0, float e1, float x) {
    const float t = saturate((x - e0) / (e1 - e0));
    return t * t * (3.f - 2.f * t);
}
FHP_HD bool analyticSample(const SceneDev& s, const Shading& sd, const Material& m, float u, f3& org, f3& dir, float& dist,
                           f3& contrib) {
    const uint32_t n = s.nALights;
    const uint32_t li = min(uint32_t(u * float(n)), n - 1u);
    const float4* L = s.alights + 4 * size_t(li);
    const float4 q0 = L[0], q1 = L[1], q2 = L[2], q3 = L[3];
    const f3 dirW = mk3(q2.x, q2.y, q2.z);
    f3 Li;
    if (f2u(q0.x) == kLightDirectional) {
        dir = -dirW;  // sampleDirectionalLight (LightHelpers.slang:190-211)
        dist = kFltMax;
        Li = mk3(q0.y, q0.z, q0.w);
    } else {
        // samplePointLight (LightHelpers.slang:217-258), spot cone and penumbra.
        const f3 toL = mk3(q1.x, q1.y, q1.z) - sd.posW;
        const float d2 = fmaxf_(dot(toL, toL), 1e-9f);
        dist = sqrtf(d2);
        dir = toL / dist;
        const float cosT = -dot(dir, dirW);
        float falloff = 1.f;
        if (cosT < q2.w) falloff = 0.f;
        else if (q3.x > 0.f) falloff = smoothstep01(0.f, q3.x, q1.w - acosf(cosT));
        Li = mk3(q0.y, q0.z, q0.w) * falloff / d2;
    }
    // Reject directions outside every lobe (getLobeTypes: a reflection lobe always exists).
    const bool hasT = m.type == MAT_STANDARD && (m.specTrans > 0.f || (m.diffTrans > 0.f && m.specTrans < 1.f && anyPos(m.diffuse)));
    const float NdotL = dot(sd.orientedFaceN(), dir);
    if (NdotL <= kMinCosTheta && !hasT) return false;
    org = rayOriginOffset(sd.posW, dot(sd.faceN, dir) >= 0.f ? sd.faceN : -sd.faceN);
    contrib = matEval(sd, m, dir) * Li * float(n);
    return true;
}

// EmissiveLightSamplerHelpers.slang:106-124
FHP_HD float triPdf(const SceneDev& s, f3 posW, uint32_t li, f3 hitPos, f3 hitN) {
    f3 toL = hitPos - posW;
    const float d2 = dot(toL, toL);
    if (d2 <= kFltMin) return 0.f;
    f3 L = toL / sqrtf(d2);
    float cosT = dot(hitN, -L);
    if (cosT <= 0.f) return 0.f;
    float denom = fmaxf_(kFltMin, cosT * s.lights[5 * size_t(li)].w);
    return d2 / denom;
}

// ------------------------------------------------------------------ packed light-BVH nodes
// The reference's PackedNode (LightBVHTypes.slang:97-198), 32 bytes: word 0 the right child's
// index (internal node, bit 31 clear) or bit 31 | triangle count << 27 | first triangle (leaf);
// words 1-3 the box centre; word 4 the half extents x, y as IEEE half floats; word 5 the half
// extent z (low half) and the cosine of the normal cone's spread quantised to 16 bits,
// (uint)((cos + 1) * 32767) rounding towards a wider cone; word 6 the cone axis in the octahedral
// map as two 16-bit snorms (encodeNormal2x16, PackedFormats.slang:53-66); word 7 the flux.
// Sampling uses the decoded values, as the reference does.
constexpr uint32_t kLNodeCountShift = 27, kLNodeCountMask = 0xfu, kLNodeOffsetMask = (1u << 27) - 1u;
struct LNodeAttr {
    f3 origin, extent, coneDir;
    float cosCone, flux;
};
// IEEE 754 binary16 with round to nearest even (overflow to infinity, NaN kept quiet).
FHP_HD uint32_t f32ToF16(float f) {
    const uint32_t x = f2u(f), sign = (x >> 16) & 0x8000u;
    const uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) return sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u);
    if (ax >= 0x477ff000u) return sign | 0x7c00u;  // rounds to beyond the largest half (65504)
    if (ax < 0x38800000u) {
        // Subnormal half (or zero): the value in units of 2^-24, rounded to nearest even.
        if (ax < 0x33000000u) return sign;  // below half of the smallest subnormal
        const uint32_t m = (ax & 0x7fffffu) | 0x800000u;
        const uint32_t shift = 126u - (ax >> 23);  // in units of 2^-24 the value is m * 2^(e - 126)
        const uint32_t q = m >> shift, rem = m & ((1u << shift) - 1u), half = 1u << (shift - 1u);
        return sign | (q + ((rem > half || (rem == half && (q & 1u))) ? 1u : 0u));
    }
    const uint32_t q = ((ax >> 13) - ((127u - 15u) << 10)), rem = ax & 0x1fffu;
    return sign | (q + ((rem > 0x1000u || (rem == 0x1000u && (q & 1u))) ? 1u : 0u));
}
FHP_HD float f16ToF32(uint32_t h) {
#if defined(__HIP_DEVICE_COMPILE__)
    // The hardware conversion (exact, like the bit-field form below).
    return float(__builtin_bit_cast(_Float16, uint16_t(h)));
#endif
    const uint32_t sign = (h & 0x8000u) << 16, e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    if (e == 0u) {
        const float v = float(m) * (1.f / 16777216.f);  // m * 2^-24, exact
        return sign ? -v : v;
    }
    if (e == 31u) return u2f(sign | 0x7f800000u | (m << 13));
    return u2f(sign | ((e + 112u) << 23) | (m << 13));
}
// floatToSnorm16 / unpackSnorm16 (FormatConversion.slang:235-251).
FHP_HD int32_t floatToSnorm16(float v) {
    v = v != v ? 0.f : fminf(fmaxf(v, -1.f), 1.f);
    return int32_t(truncf(v * 32767.f + (v >= 0.f ? 0.5f : -0.5f)));
}
FHP_HD float snorm16ToFloat(uint32_t bits16) {
    const int32_t b = int32_t(bits16 << 16) >> 16;
    return fmaxf(float(b) / 32767.f, -1.f);
}
// ndir_to_oct_snorm / oct_to_ndir_snorm (MathHelpers.slang:156-194) with 16-bit snorms.
FHP_HD uint32_t encodeOct16(f3 n) {
    const float inv = 1.f / (fabsf(n.x) + fabsf(n.y) + fabsf(n.z));
    float px = n.x * inv, py = n.y * inv;
    if (n.z < 0.f) {
        const float wx = (1.f - fabsf(py)) * (px >= 0.f ? 1.f : -1.f), wy = (1.f - fabsf(px)) * (py >= 0.f ? 1.f : -1.f);
        px = wx;
        py = wy;
    }
    return (uint32_t(floatToSnorm16(px)) & 0xffffu) | (uint32_t(floatToSnorm16(py)) << 16);
}
FHP_HD f3 decodeOct16(uint32_t w) {
    float px = snorm16ToFloat(w & 0xffffu), py = snorm16ToFloat(w >> 16);
    const float z = 1.f - fabsf(px) - fabsf(py);
    if (z < 0.f) {
        const float wx = (1.f - fabsf(py)) * (px >= 0.f ? 1.f : -1.f), wy = (1.f - fabsf(px)) * (py >= 0.f ? 1.f : -1.f);
        px = wx;
        py = wy;
    }
    return normalize(mk3(px, py, z));
}
FHP_HD void packLightNode(uint32_t w[8], uint32_t info, const LNodeAttr& a) {
    w[0] = info;
    w[1] = f2u(a.origin.x);
    w[2] = f2u(a.origin.y);
    w[3] = f2u(a.origin.z);
    w[4] = f32ToF16(a.extent.x) | (f32ToF16(a.extent.y) << 16);
    w[5] = f32ToF16(a.extent.z) | (uint32_t((a.cosCone + 1.f) * 32767.f) << 16);
    w[6] = encodeOct16(a.coneDir);
    w[7] = f2u(a.flux);
}
FHP_HD LNodeAttr unpackLightNode(uint4 q0, uint4 q1) {
    LNodeAttr a;
    a.origin = mk3(u2f(q0.y), u2f(q0.z), u2f(q0.w));
    a.extent = mk3(f16ToF32(q1.x & 0xffffu), f16ToF32(q1.x >> 16), f16ToF32(q1.y & 0xffffu));
    a.cosCone = float(q1.y >> 16) * (1.f / 32767.f) - 1.f;
    a.coneDir = decodeOct16(q1.z);
    a.flux = u2f(q1.w);
    return a;
}

// Sampling record of a light-BVH node: its packed node decoded (unpackLightNode) together with
// the importance terms that do not depend on the shading point, so the sampler does not decode
// half floats and the octahedral axis at every visit (3 x float4, kept beside the packed nodes and
// rewritten whenever they are):
//   (origin, flux), (cone axis, cos cone), (|extent|^2, max extent, sin cone, packed word 0)
constexpr uint32_t kLNodeSampleWords = 12;
FHP_HD void lightNodeSampleRecord(float w[12], uint32_t word0, const LNodeAttr& a) {
    w[0] = a.origin.x; w[1] = a.origin.y; w[2] = a.origin.z; w[3] = a.flux;
    w[4] = a.coneDir.x; w[5] = a.coneDir.y; w[6] = a.coneDir.z; w[7] = a.cosCone;
    w[8] = dot(a.extent, a.extent);
    w[9] = fmaxf_(a.extent.x, fmaxf_(a.extent.y, a.extent.z));
    w[10] = sqrtf(fmaxf_(0.f, 1.f - a.cosCone * a.cosCone));
    w[11] = u2f(word0);
}

// Light BVH node accessors: info = (right child | leaf bit, triOffset, triCount, 0).
FHP_HD float cosSubClamped(float sA, float cA, float sB, float cB) { return cA > cB ? 1.f : cA * cB + sA * sB; }
FHP_HD float sinSubClamped(float sA, float cA, float sB, float cB) { return cA > cB ? 0.f : sA * cB - cA * sB; }
// LightBVHSampler.slang:219-273 (sphere solid-angle bound, bounding and lighting cones)
FHP_HD float lbvhImportance(const SceneDev& s, f3 posW, f3 n, uint32_t node) {
    // The node's sampling record (the packed node decoded). The reference's terms with shared
    // subexpressions and hardware square roots and reciprocals: normalize(origin - posW) and
    // (origin - posW) / distance are one vector, the bounding sphere's sine sqrt(1 - cos^2) is its
    // sqrt(r^2 / d^2).
    const float4 r0 = s.lbvhS[3 * size_t(node)], r1 = s.lbvhS[3 * size_t(node) + 1], r2 = s.lbvhS[3 * size_t(node) + 2];
    const f3 cv = mk3(r0.x - posW.x, r0.y - posW.y, r0.z - posW.z);
    const float flux = r0.w, cosCone = r1.w, sqrR = r2.x, halfR = r2.y;
    const float cd2 = dot(cv, cv);
    float distance = hwSqrt(cd2);
    const f3 L = cv * hwRcp(distance);
    // boundCosineTerm (LightBVHSampler.slang:181-210) with GeometryHelpers.slang:222-240.
    float sinBC = 0.f, cosBC = -1.f;
    if (!(cd2 < sqrR)) {
        const float s2 = sqrR * hwRcp(cd2);
        cosBC = hwSqrt(1.f - s2);
        sinBC = hwSqrt(s2);
    }
    const float cosL = fminf_(fmaxf_(dot(n, L), -1.f), 1.f);
    const float sinL = hwSqrt(1.f - cosL * cosL);
    const float NdotL = saturate(cosSubClamped(sinL, cosL, sinBC, cosBC));
    if (NdotL == 0.f) return 0.f;
    float orient = 1.f;
    if (cosCone > 0.f) {
        const float sinCone = r2.z;
        const float cosT = -dot(mk3(r1.x, r1.y, r1.z), L);
        const float sinT = hwSqrt(fmaxf_(0.f, 1.f - cosT * cosT));
        const float c0 = cosSubClamped(sinT, cosT, sinCone, cosCone);
        const float s0 = sinSubClamped(sinT, cosT, sinCone, cosCone);
        orient = fmaxf_(0.f, cosSubClamped(s0, c0, sinBC, cosBC));
    }
    distance = fmaxf_(halfR, distance);
    return (flux * NdotL) * orient * hwRcp(distance * distance);
}
FHP_HD uint4 lbvhInfo(const SceneDev& s, uint32_t node) {
    const uint32_t w = f2u(s.lbvhS[3 * size_t(node) + 2].w);
    if (!(w & kLeafBit)) return make_uint4(w, 0u, 0u, 0u);
    return make_uint4(kLeafBit, w & kLNodeOffsetMask, (w >> kLNodeCountShift) & kLNodeCountMask, 0u);
}
// LightBVHSampler.slang:284-323, :375-419 (uniform pick in leaf), :430-446
FHP_HD bool lbvhSample(const SceneDev& s, f3 posW, f3 n, float u, float& pdf, uint32_t& light) {
    if (s.nLbvh == 0) return false;
    pdf = 1.f;
    uint32_t node = 0;
    uint4 info = lbvhInfo(s, node);
    while (!(info.x & kLeafBit)) {
        uint32_t l = node + 1, r = info.x;
        float li = lbvhImportance(s, posW, n, l), ri = lbvhImportance(s, posW, n, r);
        float tot = li + ri;
        if (tot == 0.f) return false;
        // pL exact (the pdf replays it); the random number is rescaled to the chosen child with a
        // hardware reciprocal, which moves the sample by about an ulp, not the pdf.
        float pL = li / tot, pR = 1.0f - pL;
        if (u < pL) { u = u * hwRcp(pL); pdf *= pL; node = l; }
        else { u = (u - pL) * hwRcp(pR); pdf *= pR; node = r; }
        info = lbvhInfo(s, node);
    }
    uint32_t cnt = info.z;
    uint32_t idx = min(uint32_t(u * float(cnt)), cnt - 1u);
    light = s.lbvhTri[info.y + idx];
    pdf *= 1.0f / float(cnt);
    return true;
}
// LightBVHSampler.slang:108-124, :455-492, :502-533
FHP_HD float lbvhSelPdf(const SceneDev& s, f3 posW, f3 n, uint32_t light) {
    if (s.nLbvh == 0) return 0.f;
    uint2 m = s.lbvhMask[light];
    uint64_t mask = (uint64_t(m.y) << 32) | m.x;
    float p = 1.f;
    uint32_t node = 0;
    uint4 info = lbvhInfo(s, node);
    while (!(info.x & kLeafBit)) {
        uint32_t l = node + 1, r = info.x;
        float li = lbvhImportance(s, posW, n, l), ri = lbvhImportance(s, posW, n, r);
        float tot = li + ri;
        if (tot == 0.f) return 0.f;
        float pL = li / tot, pR = 1.0f - pL;
        if ((mask & 1ull) == 0) { p *= pL; node = l; } else { p *= pR; node = r; }
        mask >>= 1;
        info = lbvhInfo(s, node);
    }
    if (p == 0.f) return 0.f;
    return p * (1.0f / float(info.z));
}
enum { LS_UNIFORM = 0, LS_LIGHTBVH = 1 };
// EmissiveUniformSampler.slang:55-74 / LightBVHSampler.slang:75-94
FHP_HD bool sampleLight(const SceneDev& s, int kind, f3 posW, f3 n, Rng& rng, LightSample& ls) {
    if (s.nActive == 0) return false;
    float uL = rng.u1();
    uint32_t li;
    float sel;
    if (kind == LS_UNIFORM) {
        uint32_t idx = min(uint32_t(uL * float(s.nActive)), s.nActive - 1u);
        li = s.activeLights[idx];
        sel = 1.f / float(s.nActive);
    } else {
        if (!lbvhSample(s, posW, n, uL, sel, li)) return false;
    }
    float ux = rng.u1(), uy = rng.u1();
    if (!sampleTri(s, posW, li, ux, uy, ls)) return false;
    ls.pdf *= sel;
    return true;
}
FHP_HD float lightPdf(const SceneDev& s, int kind, f3 posW, f3 n, uint32_t li, f3 hitPos, f3 hitN) {
    if (li == kInvalid) return 0.f;
    float sel = kind == LS_UNIFORM ? (s.nActive ? 1.f / float(s.nActive) : 0.f) : lbvhSelPdf(s, posW, n, li);
    return sel * triPdf(s, posW, li, hitPos, hitN);
}
// PTUtils.slang:87-93
FHP_HD float misPower(float p0, float p1) {
    float q0 = (1.f * p0) * (1.f * p0);
    float q1 = (1.f * p1) * (1.f * p1);
    return q0 / (q0 + q1);
}

}  // namespace fhp
