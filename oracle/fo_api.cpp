// C entry points of the oracle library (loaded with ctypes by tests and bench.py's CPU leg).
// TEST INFRASTRUCTURE ONLY (see fo_core.h).
#include "fo_scene.h"
#include "fo_f64.h"
#include <atomic>
#include <thread>

using namespace fo;

namespace {
unsigned hwThreads(int requested) {
    if (requested > 0) return unsigned(requested);
    unsigned n = std::thread::hardware_concurrency();
    return n ? n : 1;
}
// Runs fn(row) for rows [0, H) on `threads` workers (dynamic row scheduling).
template <typename F> void parallelRows(uint32_t H, unsigned threads, F fn) {
    std::atomic<uint32_t> next{0};
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < threads; t++)
        pool.emplace_back([&] {
            for (uint32_t r; (r = next.fetch_add(1)) < H;) fn(r);
        });
    for (auto& th : pool) th.join();
}
}  // namespace

extern "C" {

int fo_version() { return 1; }

// ---------------------------------------------------------------- PRNG / bit tricks
void fo_splitmix64(uint64_t state, int n, uint64_t* out) {
    for (int i = 0; i < n; i++) out[i] = splitmix64_next(state);
}
void fo_xoshiro128ss(const uint32_t* seed, int n, uint32_t* out) {
    SampleGen g;
    for (int i = 0; i < 4; i++) g.s[i] = seed[i];
    for (int i = 0; i < n; i++) out[i] = g.next();
}
void fo_sample_gen(uint32_t px, uint32_t py, uint32_t sampleNumber, int n, float* out) {
    SampleGen g = makeSampleGen(px, py, sampleNumber);
    for (int i = 0; i < n; i++) out[i] = g.next1D();
}
void fo_interleave(int n, const uint32_t* x, const uint32_t* y, uint32_t* out) {
    for (int i = 0; i < n; i++) out[i] = interleave_32bit(x[i], y[i]);
}
void fo_jenkins(int n, const uint32_t* in, uint32_t* out) {
    for (int i = 0; i < n; i++) out[i] = jenkinsHash(in[i]);
}
void fo_offset_ray(int n, const float* pos, const float* nrm, float* out) {
    for (int i = 0; i < n; i++) {
        V3f r = computeRayOrigin(V3f(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]), V3f(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]));
        out[3 * i] = r.x; out[3 * i + 1] = r.y; out[3 * i + 2] = r.z;
    }
}
// Watertight test of ray i against triangle i: out[4*i] = hit flag, t, b1, b2.
void fo_intersect(int n, const float* rays, const float* tris, float* out) {
    for (int i = 0; i < n; i++) {
        const float* r = rays + 6 * i;
        const float* t = tris + 9 * i;
        V3f v[3] = {V3f(t[0], t[1], t[2]), V3f(t[3], t[4], t[5]), V3f(t[6], t[7], t[8])};
        V3f b; float tt = 0.f;
        bool h = intersectRayTriangle(V3f(r[0], r[1], r[2]), V3f(r[3], r[4], r[5]), v, b, tt);
        out[4 * i] = h ? 1.f : 0.f; out[4 * i + 1] = tt; out[4 * i + 2] = b.y; out[4 * i + 3] = b.z;
    }
}
void fo_intersect_f64(int n, const double* rays, const double* tris, double* out) {
    for (int i = 0; i < n; i++) {
        const double* r = rays + 6 * i;
        const double* t = tris + 9 * i;
        V3<double> v[3] = {V3<double>(t[0], t[1], t[2]), V3<double>(t[3], t[4], t[5]), V3<double>(t[6], t[7], t[8])};
        V3<double> b; double tt = 0.;
        bool h = intersectRayTriangle(V3<double>(r[0], r[1], r[2]), V3<double>(r[3], r[4], r[5]), v, b, tt);
        out[4 * i] = h ? 1. : 0.; out[4 * i + 1] = tt; out[4 * i + 2] = b.y; out[4 * i + 3] = b.z;
    }
}
// Cosine-hemisphere, GGX VNDF and vMF warps (for statistical tests).
void fo_sample_cosine(int n, const float* u, float* out) {
    for (int i = 0; i < n; i++) {
        float pdf;
        V3f w = sample_cosine_hemisphere_concentric(u[2 * i], u[2 * i + 1], pdf);
        out[4 * i] = w.x; out[4 * i + 1] = w.y; out[4 * i + 2] = w.z; out[4 * i + 3] = pdf;
    }
}
void fo_sample_vmf(int n, const float* u, float kappa, float* out) {
    for (int i = 0; i < n; i++) {
        V3f w = sampleVonMisesFisher(u[2 * i], u[2 * i + 1], kappa);
        out[3 * i] = w.x; out[3 * i + 1] = w.y; out[3 * i + 2] = w.z;
    }
}
// Standard BSDF in the local frame. mat = [type, base rgb, roughness, metallic, ior, eta,
// specularTransmission, diffuseTransmission, transmission albedo rgb] (13 floats; the albedo is
// used as given, after any thin-surface square root).
static BsdfData<float> bsdfParams(const float* mat) {
    BsdfData<float> d;
    d.type = int(mat[0]);
    V3f base(mat[1], mat[2], mat[3]);
    float ior = mat[6], f = (ior - 1.f) / (ior + 1.f), F0 = f * f, metal = mat[5];
    d.diffuse = base + (V3f(0.f) - base) * metal;
    d.specular = V3f(F0) + (base - V3f(F0)) * metal;
    d.roughness = mat[4];
    d.metallic = metal;
    d.eta = mat[7];
    d.specTrans = mat[8];
    d.diffTrans = mat[9];
    d.transmission = V3f(mat[10], mat[11], mat[12]);
    return d;
}
// eval: out[4*i] = f rgb, pdf.
void fo_bsdf_eval(int n, const float* mat, const float* wi, const float* wo, float* out) {
    BsdfData<float> d = bsdfParams(mat);
    for (int i = 0; i < n; i++) {
        V3f a(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]), b(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]);
        StdBsdf<float> bs;
        bs.init(a, d);
        V3f v = bs.eval(a, b);
        out[4 * i] = v.x; out[4 * i + 1] = v.y; out[4 * i + 2] = v.z; out[4 * i + 3] = bs.evalPdf(a, b);
    }
}
// sample: out[8*i] = wo xyz, weight rgb, pdf, valid (delta samples report valid = 2).
void fo_bsdf_sample(int n, const float* mat, const float* wi, const uint32_t* seeds, float* out) {
    BsdfData<float> d = bsdfParams(mat);
    for (int i = 0; i < n; i++) {
        V3f a(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]);
        SampleGen g = makeSampleGen(seeds[2 * i], 0, seeds[2 * i + 1]);
        StdBsdf<float> bs;
        bs.init(a, d);
        V3f wo, w; float pdf; bool delta;
        bool ok = bs.sample(a, wo, pdf, w, delta, g);
        float* o = out + 8 * i;
        o[0] = wo.x; o[1] = wo.y; o[2] = wo.z; o[3] = w.x; o[4] = w.y; o[5] = w.z; o[6] = pdf;
        o[7] = ok ? (delta ? 2.f : 1.f) : 0.f;
    }
}

// Dielectric Fresnel: out[2*i] = F, cosT.
void fo_fresnel_dielectric(int n, const float* eta, const float* cosI, float* out) {
    for (int i = 0; i < n; i++) {
        float ct;
        out[2 * i] = fresnelDielectric(eta[i], cosI[i], ct);
        out[2 * i + 1] = ct;
    }
}

// ---------------------------------------------------------------- scene
// cam = [pos xyz, target xyz, up xyz, fovY(rad), aspect, focalDistance, nearZ, farZ]
void* fo_scene_create(uint32_t nV, const float* pos, const float* nrm, const float* uv, uint32_t nT, const uint32_t* idx,
                      const uint32_t* triMat, uint32_t nM, const float* mats, uint32_t nTex, const int32_t* texDesc,
                      uint64_t nTexels, const float* texels, const float* cam) {
    Scene* s = new Scene();
    s->nV = nV; s->nT = nT; s->nM = nM; s->nTex = nTex;
    s->pos.resize(nV); s->nrm.resize(nV);
    for (uint32_t i = 0; i < nV; i++) {
        s->pos[i] = V3f(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        s->nrm[i] = V3f(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
    }
    s->uv.assign(uv, uv + 2 * size_t(nV));
    s->idx.assign(idx, idx + 3 * size_t(nT));
    s->triMat.assign(triMat, triMat + nT);
    s->mats.assign(mats, mats + size_t(nM) * M_STRIDE);
    if (nTex) {
        s->texDesc.assign(texDesc, texDesc + 4 * size_t(nTex));
        s->texels.assign(texels, texels + nTexels);
    }
    s->cam = makeCamera(V3f(cam[0], cam[1], cam[2]), V3f(cam[3], cam[4], cam[5]), V3f(cam[6], cam[7], cam[8]), cam[9], cam[10],
                        cam[11]);
    s->cam.nearZ = cam[12];
    s->cam.farZ = cam[13];
    buildOracleBVH(*s);
    buildLightCollection(*s);
    buildLightBVH(*s);
    return s;
}
void fo_scene_destroy(void* h) { delete static_cast<Scene*>(h); }
// Animation update: new vertex positions and normals for every vertex (3 floats each), then
// the traversal structure, the emissive triangles (positions, normals, areas; flux kept) and a
// refit of the light BVH, as Scene::update does after a node transform.
void fo_scene_animate(void* h, const float* pos, const float* nrm) {
    Scene* s = static_cast<Scene*>(h);
    for (uint32_t i = 0; i < s->nV; i++) {
        s->pos[i] = V3f(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        if (nrm) s->nrm[i] = V3f(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
    }
    buildOracleBVH(*s);
    updateLightPositions(*s);
    refitLightBVH(*s);
}
// Analytic lights: records [type, intensity rgb, position, openingAngle, direction,
// penumbraAngle, ...] (16 floats); Light::setWorldDirection normalises, PointLight's angle
// setters clamp (Light.cpp:174-223).
void fo_scene_set_analytic_lights(void* h, uint32_t n, const float* recs) {
    Scene* s = static_cast<Scene*>(h);
    s->aLights.assign(16 * size_t(n), 0.f);
    for (uint32_t i = 0; i < n; i++) {
        const float* q = recs + 16 * size_t(i);
        float* r = &s->aLights[16 * size_t(i)];
        const float pi = Consts<float>::pi;
        float opening = std::min(std::max(q[7], 0.f), pi);
        float penumbra = std::min(std::max(q[11], 0.f), opening);
        V3f d = normalize(V3f(q[8], q[9], q[10]));
        r[0] = q[0]; r[1] = q[1]; r[2] = q[2]; r[3] = q[3];
        r[4] = q[4]; r[5] = q[5]; r[6] = q[6]; r[7] = opening;
        r[8] = d.x; r[9] = d.y; r[10] = d.z; r[11] = std::cos(opening); r[12] = penumbra;
    }
}
// Environment map: lat-long texels H x W x 3, params [intensity, tint rgb, rotation degrees xyz],
// flags 1 light / 2 background; builds EnvMapSampler's importance map (dim x dim, ns x ns samples
// per texel, EnvMapSampler.cpp:39-115). W = 0 removes the map.
void fo_scene_set_env_map(void* h, const float* texels, uint32_t W, uint32_t H, const float* prm, uint32_t flags,
                          uint32_t dim, uint32_t ns, int threads) {
    Scene* s = static_cast<Scene*>(h);
    s->envImp.clear();
    s->envImpDim = 0;
    if (W == 0 || H == 0) { s->envTex.clear(); s->envW = s->envH = s->envFlags = 0; return; }
    s->envTex.assign(texels, texels + 3 * size_t(W) * H);
    s->envW = W; s->envH = H; s->envFlags = flags;
    // EnvMap::setRotation (EnvMap.cpp:62-72): matrixFromRotationXYZ of the negated angles
    // (MatrixMath.h:649-679); world-to-local is its transpose.
    const float k = Consts<float>::pi / 180.f;
    const float c1 = std::cos(-prm[4] * k), c2 = std::cos(-prm[5] * k), c3 = std::cos(-prm[6] * k);
    const float s1 = std::sin(-prm[4] * k), s2 = std::sin(-prm[5] * k), s3 = std::sin(-prm[6] * k);
    const float m[3][3] = {{c2 * c3, c2 * s3, -s2},
                           {-c1 * s3 + s1 * s2 * c3, c1 * c3 + s1 * s2 * s3, s1 * c2},
                           {s1 * s3 + c1 * s2 * c3, -s1 * c3 + c1 * s2 * s3, c1 * c2}};
    for (int r = 0; r < 3; r++) s->envInv[r] = V3f(m[0][r], m[1][r], m[2][r]);
    s->envScale = V3f(prm[0] * prm[1], prm[0] * prm[2], prm[0] * prm[3]);
    std::vector<float> imp((4 * size_t(dim) * dim - 1) / 3);
    float* fine = &imp[(size_t(dim) * dim - 1) / 3];
    parallelRows(dim, hwThreads(threads), [&](uint32_t y) {
        for (uint32_t x = 0; x < dim; x++) fine[size_t(y) * dim + x] = envImportanceTexel(*s, x, y, dim, ns);
    });
    for (uint32_t r = dim / 2; r >= 1; r /= 2) {
        const float* src = &imp[(4 * size_t(r) * r - 1) / 3];
        float* dst = &imp[(size_t(r) * r - 1) / 3];
        const uint32_t R = 2 * r;
        for (uint32_t y = 0; y < r; y++)
            for (uint32_t x = 0; x < r; x++) {
                const float* q = src + size_t(2 * y) * R + 2 * x;
                dst[size_t(y) * r + x] = ((q[0] + q[1]) + (q[R] + q[R + 1])) * 0.25f;  // generateMips: 2 x 2 mean
            }
    }
    if (imp[0] > 0.f && std::isfinite(imp[0])) { s->envImp.swap(imp); s->envImpDim = dim; }
}
uint32_t fo_env_importance(void* h, float* out) {
    Scene* s = static_cast<Scene*>(h);
    if (out) std::copy(s->envImp.begin(), s->envImp.end(), out);
    return s->envImpDim;
}
// EnvMapSampler::sample for n pairs u: out 8 floats per sample (dir, pdf, Le, valid).
void fo_env_sample(void* h, const float* u, uint32_t n, float* out) {
    Scene* s = static_cast<Scene*>(h);
    for (uint32_t i = 0; i < n; i++) {
        V3f d, Le; float pdf;
        bool ok = envSample(*s, u[2 * i], u[2 * i + 1], d, pdf, Le);
        float* o = out + 8 * size_t(i);
        o[0] = d.x; o[1] = d.y; o[2] = d.z; o[3] = pdf; o[4] = Le.x; o[5] = Le.y; o[6] = Le.z; o[7] = ok ? 1.f : 0.f;
    }
}
void fo_env_pdf(void* h, const float* dirs, uint32_t n, float* out) {
    Scene* s = static_cast<Scene*>(h);
    for (uint32_t i = 0; i < n; i++) out[i] = envPdf(*s, V3f(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
}
void fo_env_eval(void* h, const float* dirs, uint32_t n, float* out) {
    Scene* s = static_cast<Scene*>(h);
    for (uint32_t i = 0; i < n; i++) {
        V3f L = envEval(*s, V3f(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
        out[3 * i] = L.x; out[3 * i + 1] = L.y; out[3 * i + 2] = L.z;
    }
}
// Equal-area octahedral map and its inverse for n points uv: dirs (3 floats), uvBack (2 floats).
void fo_oct_equal_area(const float* uv, uint32_t n, float* dirs, float* uvBack) {
    for (uint32_t i = 0; i < n; i++) {
        V3f d = octToDirEqualArea(uv[2 * i], uv[2 * i + 1]);
        dirs[3 * i] = d.x; dirs[3 * i + 1] = d.y; dirs[3 * i + 2] = d.z;
        dirToOctEqualArea(d, uvBack[2 * i], uvBack[2 * i + 1]);
    }
}
void fo_scene_set_materials(void* h, const float* mats) {
    Scene* s = static_cast<Scene*>(h);
    s->mats.assign(mats, mats + size_t(s->nM) * M_STRIDE);
}
void fo_scene_set_texels(void* h, const float* texels) {
    Scene* s = static_cast<Scene*>(h);
    std::copy(texels, texels + s->texels.size(), s->texels.begin());
}
// Frozen-sampling mode for finite differences: sampling/pdfs use these parameters while
// evaluation uses the ones set by fo_scene_set_materials / fo_scene_set_texels. Passing null
// materials turns the mode off.
void fo_scene_set_sampling_params(void* h, const float* mats, const float* texels) {
    Scene* s = static_cast<Scene*>(h);
    if (!mats) { s->frozen = false; s->matsS.clear(); s->texelsS.clear(); return; }
    s->frozen = true;
    s->matsS.assign(mats, mats + size_t(s->nM) * M_STRIDE);
    s->texelsS = texels ? std::vector<float>(texels, texels + s->texels.size()) : s->texels;
}
// Camera basis as computed on the host: out = [U, V, W] (9 floats).
void fo_scene_camera(void* h, float* out) {
    Scene* s = static_cast<Scene*>(h);
    V3f v[3] = {s->cam.U, s->cam.V, s->cam.W};
    for (int i = 0; i < 3; i++) { out[3 * i] = v[i].x; out[3 * i + 1] = v[i].y; out[3 * i + 2] = v[i].z; }
}
uint32_t fo_scene_light_count(void* h) { return uint32_t(static_cast<Scene*>(h)->lights.size()); }
uint32_t fo_scene_active_light_count(void* h) { return uint32_t(static_cast<Scene*>(h)->activeLights.size()); }
// Light data: per light [p0 p1 p2 normal area Le flux tri] = 9+3+1+3+1+1 = 18 floats.
void fo_scene_lights(void* h, float* out) {
    Scene* s = static_cast<Scene*>(h);
    for (size_t i = 0; i < s->lights.size(); i++) {
        const EmissiveTri& e = s->lights[i];
        float* o = out + 18 * i;
        for (int j = 0; j < 3; j++) { o[3 * j] = e.p[j].x; o[3 * j + 1] = e.p[j].y; o[3 * j + 2] = e.p[j].z; }
        o[9] = e.normal.x; o[10] = e.normal.y; o[11] = e.normal.z; o[12] = e.area;
        o[13] = e.Le.x; o[14] = e.Le.y; o[15] = e.Le.z; o[16] = e.flux; o[17] = asfloat(e.tri);
    }
}
uint32_t fo_lbvh_node_count(void* h) { return uint32_t(static_cast<Scene*>(h)->lnodes.size()); }
// Node: [origin xyz, extent xyz, flux, coneDir xyz, cosCone, leaf, right, triOffset, triCount] (15 floats;
// integer fields stored as float values).
void fo_lbvh_nodes(void* h, float* out, uint32_t* triIdx, uint64_t* bitmasks) {
    Scene* s = static_cast<Scene*>(h);
    for (size_t i = 0; i < s->lnodes.size(); i++) {
        const Scene::LNode& n = s->lnodes[i];
        float* o = out + 15 * i;
        o[0] = n.origin.x; o[1] = n.origin.y; o[2] = n.origin.z; o[3] = n.extent.x; o[4] = n.extent.y; o[5] = n.extent.z;
        o[6] = n.flux; o[7] = n.coneDir.x; o[8] = n.coneDir.y; o[9] = n.coneDir.z; o[10] = n.cosCone;
        o[11] = n.leaf ? 1.f : 0.f; o[12] = float(n.right); o[13] = float(n.triOffset); o[14] = float(n.triCount);
    }
    std::copy(s->lTriIdx.begin(), s->lTriIdx.end(), triIdx);
    std::copy(s->lBitmask.begin(), s->lBitmask.end(), bitmasks);
}
// Light sampling at n shading points (seeds: 2 words per point). out per point:
// [selected light (-1 if selection failed), selection pdf, full solid-angle pdf (0 if the
// triangle sample was rejected), 1 if the whole sample is valid].
void fo_light_sample(void* h, int kind, int n, const float* pos, const float* nrm, const uint32_t* seeds, float* out) {
    Scene* s = static_cast<Scene*>(h);
    for (int i = 0; i < n; i++) {
        V3f p(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]), nn(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
        SampleGen g = makeSampleGen(seeds[2 * i], 0, seeds[2 * i + 1]);
        SampleGen g2 = g;
        LightSample ls;
        bool ok = sampleLight(*s, kind, p, nn, g, ls);
        // Replay the selection step alone (same generator state) to report it even on rejection.
        float u = g2.next1D();
        uint32_t li = kInvalid;
        float sel = 0.f;
        bool selOk = false;
        if (!s->activeLights.empty()) {
            if (kind == LS_UNIFORM) {
                uint32_t cnt = uint32_t(s->activeLights.size());
                li = s->activeLights[std::min(uint32_t(u * float(cnt)), cnt - 1)];
                sel = 1.f / float(cnt);
                selOk = true;
            } else {
                selOk = lbvhSample(*s, p, nn, u, sel, li);
            }
        }
        out[4 * i] = selOk ? float(li) : -1.f;
        out[4 * i + 1] = selOk ? sel : 0.f;
        out[4 * i + 2] = ok ? ls.pdf : 0.f;
        out[4 * i + 3] = ok ? 1.f : 0.f;
    }
}
// Selection pdf of every active light from one shading point (LightBVH traversal pdf).
void fo_light_selection_pdfs(void* h, int kind, const float* pos, const float* nrm, float* out) {
    Scene* s = static_cast<Scene*>(h);
    V3f p(pos[0], pos[1], pos[2]), nn(nrm[0], nrm[1], nrm[2]);
    for (size_t i = 0; i < s->lights.size(); i++)
        out[i] = kind == LS_UNIFORM ? (s->lights[i].flux > 0.f ? 1.f / float(s->activeLights.size()) : 0.f)
                                    : lbvhSelectionPdf(*s, p, nn, uint32_t(i));
}

// ---------------------------------------------------------------- tracing
// rays: 8 floats per ray (o xyz, tmin, d xyz, tmax); out: 4 words per ray (tri, t, b1, b2).
void fo_trace(void* h, int n, const float* rays, uint32_t* out, int brute, int anyHit, int threads) {
    Scene* s = static_cast<Scene*>(h);
    unsigned th = hwThreads(threads);
    uint32_t rows = uint32_t((n + 1023) / 1024);
    parallelRows(rows, th, [&](uint32_t r) {
        for (int i = int(r) * 1024; i < std::min(n, int(r + 1) * 1024); i++) {
            const float* q = rays + 8 * i;
            V3f o(q[0], q[1], q[2]), d(q[4], q[5], q[6]);
            float tmin = q[3], tmax = q[7];
            Hit hit;
            if (brute) {
                if (anyHit) { for (uint32_t t = 0; t < s->nT && !hit.valid(); t++) testTri(*s, t, o, d, tmin, tmax, hit); }
                else hit = bruteClosest(*s, o, d, tmin, tmax);
            } else {
                hit = traceBVH(*s, o, d, tmin, tmax, anyHit != 0);
            }
            out[4 * i] = hit.tri;
            out[4 * i + 1] = asuint(hit.t); out[4 * i + 2] = asuint(hit.u); out[4 * i + 3] = asuint(hit.v);
        }
    });
}

// ---------------------------------------------------------------- integrators
// MinimalPathTracer render: out rgb (H*W*3), spp samples per pixel with sample numbers
// frame0*spp + s, averaged like AccumulatePass Single mode (Accumulate.cs.slang:76-114).
// flags: bit0 computeDirect, bit1 useImportanceSampling, bit2 useEmissiveNEE, bit3 brute, bit4 useEnvNEE.
// stats (optional): [closest, visibility, nodes, tris].
void fo_render_mpt(void* h, uint32_t W, uint32_t H, uint32_t spp, uint32_t frame0, uint32_t maxBounces, uint32_t flags,
                   int lightSampler, int threads, float* out, uint64_t* stats) {
    Scene* s = static_cast<Scene*>(h);
    MptParams P;
    P.maxBounces = maxBounces;
    P.computeDirect = flags & 1;
    P.useImportanceSampling = flags & 2;
    P.useEmissiveNEE = flags & 4;
    P.brute = flags & 8;
    P.useEnvNEE = flags & 16;
    P.lightSampler = lightSampler;
    std::vector<TraceStats> rowStats(H);
    parallelRows(H, hwThreads(threads), [&](uint32_t y) {
        Tracer tr{*s, P.brute, stats ? &rowStats[y] : nullptr, nullptr};
        for (uint32_t x = 0; x < W; x++) {
            V3f sum(0.f);
            for (uint32_t k = 0; k < spp; k++) sum = sum + mptTracePath(*s, P, x, y, W, H, frame0 * spp + k, tr);
            V3f o = sum * (1.0f / float(spp));
            float* q = out + 3 * (size_t(y) * W + x);
            q[0] = o.x; q[1] = o.y; q[2] = o.z;
        }
    });
    if (stats) {
        TraceStats t;
        for (auto& r : rowStats) { t.closest += r.closest; t.visibility += r.visibility; t.nodes += r.nodes; t.tris += r.tris; }
        stats[0] = t.closest; stats[1] = t.visibility; stats[2] = t.nodes; stats[3] = t.tris;
    }
}
// MinimalPathTracer in double precision (fo_f64.h): per-pixel mean over samples frame0 * spp + k,
// out = H x W x 3 doubles. flags: 1 computeDirect, 2 useImportanceSampling.
void fo_render_mpt_f64(void* h, uint32_t W, uint32_t H, uint32_t spp, uint32_t frame0, uint32_t maxBounces, uint32_t flags,
                       int threads, double* out) {
    Scene* s = static_cast<Scene*>(h);
    parallelRows(H, hwThreads(threads), [&](uint32_t y) {
        for (uint32_t x = 0; x < W; x++) {
            f64::V3d sum(0.0);
            for (uint32_t k = 0; k < spp; k++)
                sum = sum + f64::tracePathD(*s, maxBounces, flags & 1, flags & 2, x, y, W, H, frame0 * spp + k);
            double* q = out + 3 * (size_t(y) * W + x);
            q[0] = sum.x / spp; q[1] = sum.y / spp; q[2] = sum.z / spp;
        }
    });
}
// Rays traced by MinimalPathTracer for pixels [p0, p1) (row-major), sample number `sampleNumber`.
// rec: per ray 13 words: o xyz, tmin, d xyz, tmax, tri, t, b1, b2, kind (0 closest, 1 visibility). Returns the ray count
// (at most `cap` are written).
uint64_t fo_record_mpt(void* h, uint32_t W, uint32_t H, uint32_t p0, uint32_t p1, uint32_t sampleNumber, uint32_t maxBounces,
                       uint32_t flags, int lightSampler, uint64_t cap, float* rec) {
    Scene* s = static_cast<Scene*>(h);
    MptParams P;
    P.maxBounces = maxBounces;
    P.computeDirect = flags & 1;
    P.useImportanceSampling = flags & 2;
    P.useEmissiveNEE = flags & 4;
    P.brute = flags & 8;
    P.useEnvNEE = flags & 16;
    P.lightSampler = lightSampler;
    std::vector<RayRecord> recs;
    Tracer tr{*s, P.brute, nullptr, &recs};
    for (uint32_t p = p0; p < p1; p++) mptTracePath(*s, P, p % W, p / W, W, H, sampleNumber, tr);
    for (size_t i = 0; i < recs.size() && i < cap; i++) {
        const RayRecord& r = recs[i];
        float* q = rec + 13 * i;
        q[0] = r.o.x; q[1] = r.o.y; q[2] = r.o.z; q[3] = r.tmin; q[4] = r.d.x; q[5] = r.d.y; q[6] = r.d.z; q[7] = r.tmax;
        q[8] = asfloat(r.tri); q[9] = r.t; q[10] = r.u; q[11] = r.v; q[12] = asfloat(r.kind);
    }
    return recs.size();
}
// WARDiffPathTracer primal (useWAR = false), one frame per sample: frameSeed = frame0 + k.
// flags: bit0 useBSDFSampling, bit1 useNEE, bit2 useMIS, bit3 useAntithetic, bit4 brute.
// WAR primal (WARDiffPathTracer.rt.slang:114-143) of the pixels [x0, x1) x [y0, y1) of a W x H frame
// (the camera and the pixel seeds are those of the full frame); out holds the rectangle row-major.
void fo_render_war_rect(void* h, uint32_t W, uint32_t H, uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1, uint32_t spp,
                        uint32_t frame0, uint32_t maxBounces, uint32_t flags, int threads, float* out, uint64_t* stats) {
    Scene* s = static_cast<Scene*>(h);
    WarParams P;
    P.maxBounces = maxBounces;
    P.useBSDFSampling = flags & 1;
    P.useNEE = flags & 2;
    P.useMIS = flags & 4;
    P.useAntithetic = flags & 8;
    P.brute = flags & 16;
    const uint32_t RW = x1 - x0, RH = y1 - y0;
    std::vector<TraceStats> rowStats(RH);
    parallelRows(RH, hwThreads(threads), [&](uint32_t r) {
        const uint32_t y = y0 + r;
        Tracer tr{*s, P.brute, stats ? &rowStats[r] : nullptr, nullptr};
        for (uint32_t x = x0; x < x1; x++) {
            V3f sum(0.f);
            for (uint32_t k = 0; k < spp; k++) sum = sum + warTracePaths(*s, P, x, y, W, H, frame0 + k, tr);
            V3f o = sum * (1.0f / float(spp));
            float* q = out + 3 * (size_t(r) * RW + (x - x0));
            q[0] = o.x; q[1] = o.y; q[2] = o.z;
        }
    });
    if (stats) {
        TraceStats t;
        for (auto& rs : rowStats) { t.closest += rs.closest; t.visibility += rs.visibility; t.nodes += rs.nodes; t.tris += rs.tris; }
        stats[0] = t.closest; stats[1] = t.visibility; stats[2] = t.nodes; stats[3] = t.tris;
    }
}

void fo_render_war(void* h, uint32_t W, uint32_t H, uint32_t spp, uint32_t frame0, uint32_t maxBounces, uint32_t flags,
                   int threads, float* out, uint64_t* stats) {
    fo_render_war_rect(h, W, H, 0, 0, W, H, spp, frame0, maxBounces, flags, threads, out, stats);
}

// Packed light-BVH node precision of n values (test hook): half-float values, quantised cone
// cosines and octahedral 2 x 16-bit directions (3 floats each).
void fo_packed_precision(uint32_t n, const float* v, float* half, float* cone, const float* dirs, float* oct) {
    for (uint32_t i = 0; i < n; i++) {
        half[i] = halfPrecision(v[i]);
        cone[i] = conePrecision(v[i]);
        const V3f d = octNormalPrecision(V3f(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
        oct[3 * i] = d.x;
        oct[3 * i + 1] = d.y;
        oct[3 * i + 2] = d.z;
    }
}
}  // extern "C"
