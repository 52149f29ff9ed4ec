// Wavefront kernels of the MinimalPathTracer forward pass and the raw ray-query kernels.
//
// Frame structure (one launch sequence per frame, fixed grid shapes so it can be captured):
//   primary rays (per pixel) -> closest hit -> shadePrimary (per path: emission, NEE, scatter)
//   -> for depth 0..maxBounces: [visibility rays of the previous shading + closest hit, hits
//   compacted] -> shade -> resolve.
// Queues are compacted with a wave ballot + one atomic per wave (per block in the shading
// kernels); per-path state lives in
// path-indexed slots so queue order never changes results.
#include "fhp_internal.h"
#include "fhp_trace.h"

namespace fhp {

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ uint32_t waveAppend(bool pred, uint32_t* counter) {
    const uint64_t mask = __ballot(pred);
    uint32_t base = 0;
    if (mask) {
        const int leader = __ffsll((unsigned long long)mask) - 1;
        if (int(__lane_id()) == leader) base = atomicAdd(counter, uint32_t(__popcll(mask)));
        base = __shfl(base, leader);
    }
    return base + uint32_t(__popcll(mask & __lanemask_lt()));
}

// Block-wide appends: one atomic per block instead of one per wave (same-address atomics from
// every CU serialise in the L2). Every thread of the block calls them the same number of times;
// lds holds T / 64 + 1 words. blockAppendN reserves cnt consecutive slots per thread, in thread
// order, and returns the first; blockAppend reserves one slot where pred holds.
template <int T>
__device__ __forceinline__ uint32_t blockAppendN(uint32_t cnt, uint32_t* counter, uint32_t* lds) {
    static_assert(T % 64 == 0 && T <= 1024, "block of whole waves");
    const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
    uint32_t x = cnt;  // inclusive scan over the wave
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d);
        if (lane >= uint32_t(d)) x += y;
    }
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int w = 0; w < T / 64; w++) {
            const uint32_t v = lds[w];
            lds[w] = tot;
            tot += v;
        }
        lds[T / 64] = tot ? atomicAdd(counter, tot) : 0u;
    }
    __syncthreads();
    const uint32_t r = lds[T / 64] + lds[wave] + x - cnt;
    __syncthreads();  // lds is reused by the next call
    return r;
}
template <int T>
__device__ __forceinline__ uint32_t blockAppend(bool pred, uint32_t* counter, uint32_t* lds) {
    static_assert(T % 64 == 0 && T <= 1024, "block of whole waves");
    const uint64_t mask = __ballot(pred);
    const uint32_t wave = threadIdx.x >> 6;
    if (__lane_id() == 0) lds[wave] = uint32_t(__popcll(mask));
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int w = 0; w < T / 64; w++) {
            const uint32_t v = lds[w];
            lds[w] = tot;
            tot += v;
        }
        lds[T / 64] = tot ? atomicAdd(counter, tot) : 0u;
    }
    __syncthreads();
    const uint32_t r = lds[T / 64] + lds[wave] + uint32_t(__popcll(mask & __lanemask_lt()));
    __syncthreads();
    return r;
}


int numCUs() {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            cus = 256;
    }
    return cus;
}

// ------------------------------------------------------------------ raw ray queries
template <int B, int D, bool ANY>
__global__ __launch_bounds__(B) void kTraceRays(SceneDev s, const float4* __restrict__ rays, uint32_t n,
                                                uint4* __restrict__ hits, uint32_t* __restrict__ occl, TraceScratch ts) {
    __shared__ uint32_t lds[D * B];
    TraceStack st = makeStack<B, D>(lds, ts, blockIdx.x, threadIdx.x);
    const uint32_t G = gridDim.x * B;
    for (uint32_t i = blockIdx.x * B + threadIdx.x; i < n; i += G) {
        const float4 a = rays[2 * i], b = rays[2 * i + 1];
        HitRec h;
        bool hit = traceBVH4<ANY, false>(s, mk3(a.x, a.y, a.z), mk3(b.x, b.y, b.z), a.w, b.w, st, h);
        if (ANY) occl[i] = hit ? 1u : 0u;
        else hits[i] = hit ? make_uint4(h.tri, f2u(h.t), f2u(h.b1), f2u(h.b2)) : make_uint4(kInvalid, 0u, 0u, 0u);
    }
}

hipError_t launchTraceRays(const SceneDev& s, const float4* rays, uint32_t n, uint4* hits, uint32_t* occluded, bool any,
                           const TraceScratch& ts, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    return withTraceConfig(ts, [&](auto Bc, auto Dc) {
        constexpr int B = decltype(Bc)::value, D = decltype(Dc)::value;
        uint32_t blocks = std::min<uint32_t>(ts.gridBlocks, (n + B - 1) / B);
        if (any) hipLaunchKernelGGL((kTraceRays<B, D, true>), dim3(blocks), dim3(B), 0, stream, s, rays, n, hits, occluded, ts);
        else hipLaunchKernelGGL((kTraceRays<B, D, false>), dim3(blocks), dim3(B), 0, stream, s, rays, n, hits, occluded, ts);
        return hipGetLastError();
    });
}

// ------------------------------------------------------------------ MinimalPathTracer wavefront
__global__ void kZeroU32(uint32_t* p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// View of a pixel or path: one frame (the kernel's scene and MptArgs) or a batch (binary search
// of the view table by first pixel / first path).
struct ViewCtx {
    const SceneDev* s;
    const CameraDev* cam;
    uint32_t W, H, spp, frame, pixBase, pathBase;
    float4* out;
};
__device__ __forceinline__ uint32_t findView(const ViewDev* v, uint32_t n, uint32_t x, bool byPath) {
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        const uint32_t b = byPath ? v[mid].pathBase : v[mid].pixBase;
        if (b <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
// The view of path x (byPath: a path with a primary hit, whose view index kCompactPrimary stored
// in a.pathView) or of pixel x (binary search).
template <bool MULTI>
__device__ __forceinline__ ViewCtx viewCtx(const SceneDev& s0, const MptArgs& a, uint32_t x, bool byPath) {
    if (!MULTI) return ViewCtx{&s0, &s0.cam, a.W, a.H, a.spp, a.counters[2 * kCounterStride], 0u, 0u, nullptr};
    const ViewDev& v = a.views[byPath ? uint32_t(a.pathView[x]) : findView(a.views, a.nViews, x, false)];
    return ViewCtx{a.scenes + v.scene, &v.cam, v.W, v.H, v.spp, v.frame, v.pixBase, v.pathBase, v.out};
}

// Pinhole primary ray through the pixel centre (Camera.slang:46-77, VBufferRT.slang:56-66).
// Camera::computeRayPinhole (Camera.slang:46-78); (jx, jy) is the camera jitter in screen units
// (jitterX, jitterY), zero without a sample pattern.
__device__ __forceinline__ void primaryRay(const CameraDev& c, uint32_t px, uint32_t py, uint32_t W, uint32_t H, f3& o, f3& d,
                                           float& tmin, float& tmax, float jx = 0.f, float jy = 0.f) {
    float p_x = (float(px) + 0.5f) / float(W), p_y = (float(py) + 0.5f) / float(H);
    p_x = p_x + -jx;
    p_y = p_y + jy;
    float ndc_x = 2.f * p_x + -1.f, ndc_y = -2.f * p_y + 1.f;
    f3 dir = c.U * ndc_x + c.V * ndc_y + c.W;
    d = normalize(dir);
    o = c.pos;
    float invCos = 1.f / dot(normalize(c.W), d);
    tmin = c.nearZ * invCos;
    tmax = c.farZ * invCos;
}

// ------------------------------------------------------------------ PackedHitInfo
// packUnorm16_unsafe (FormatConversion.slang:292-295).
__device__ __forceinline__ uint32_t packUnorm16(float v) { return uint32_t(truncf(v * 65535.f + 0.5f)); }
__device__ __forceinline__ uint32_t lowMask(uint32_t bits) { return uint32_t((1ull << bits) - 1ull); }

// TriangleHit::pack with HitInfo::packHeader (HitInfo.slang:94-121, :462-473): type Triangle (1).
__device__ __forceinline__ void packTriangleHit(const HitInfoLayout& L, uint32_t inst, uint32_t prim, float b1, float b2,
                                                uint32_t w[4]) {
    const uint32_t typeWord = 1u << (32u - L.typeBits);
    w[0] = w[1] = w[2] = w[3] = 0u;
    if (L.typeBits + L.instBits + L.primBits <= 32) {
        w[0] = typeWord | (inst << L.primBits) | prim;
    } else {
        w[0] = typeWord | inst;
        w[1] = prim;
    }
    if (L.compressed) {
        w[1] = (packUnorm16(b2) << 16) | packUnorm16(b1);
    } else {
        w[2] = f2u(b1);
        w[3] = f2u(b2);
    }
}

// TriangleHit(PackedHitInfo) with HitInfo::unpackHeader (HitInfo.slang:94-108, :475-487). Returns
// false unless the record is a triangle hit of an existing mesh and primitive with finite
// barycentrics (records that do not name a triangle of this scene read as misses).
__device__ __forceinline__ bool unpackTriangleHit(const SceneDev& s, const uint32_t* w, uint32_t& tri, float& b1, float& b2) {
    const HitInfoLayout& L = s.hit;
    if ((w[0] >> (32u - L.typeBits)) != 1u || !s.meshTriOff) return false;
    uint32_t inst, prim;
    if (L.typeBits + L.instBits + L.primBits <= 32) {
        inst = (w[0] >> L.primBits) & lowMask(L.instBits);
        prim = w[0] & lowMask(L.primBits);
    } else {
        inst = w[0] & lowMask(L.instBits);
        prim = w[1];
    }
    if (inst >= s.nMeshes) return false;
    const uint32_t t0 = s.meshTriOff[inst], t1 = s.meshTriOff[inst + 1];
    if (prim >= t1 - t0) return false;
    tri = t0 + prim;
    if (L.compressed) {
        // unpackUnorm2x16 (FormatConversion.slang:338-341)
        b1 = float(w[1] & 0xffffu) * (1.f / 65535);
        b2 = float(w[1] >> 16) * (1.f / 65535);
    } else {
        b1 = u2f(w[2]);
        b2 = u2f(w[3]);
    }
    return isfinite(b1) && isfinite(b2);
}

// Mesh (geometry instance) holding global triangle tri: the last m with meshTriOff[m] <= tri.
__device__ __forceinline__ uint32_t meshOfTriangle(const SceneDev& s, uint32_t tri) {
    uint32_t lo = 0, hi = s.nMeshes;  // answer in [lo, hi)
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s.meshTriOff[mid] <= tri) lo = mid;
        else hi = mid;
    }
    return lo;
}

// VBufferRT::writeHit / writeMiss / writeAux (VBufferRT.slang:76-143) for the pinhole camera
// without jitter (samplePattern Center). depth = NDC z / w of the interpolated hit position.
template <int B, int D>
__global__ __launch_bounds__(B) void kVBuffer(SceneDev s, uint32_t W, uint32_t H, float jx, float jy, uint32_t* __restrict__ vbuf,
                                              float* __restrict__ depth, float4* __restrict__ viewW,
                                              float* __restrict__ mask, TraceScratch ts) {
    __shared__ uint32_t lds[D * B];
    TraceStack st = makeStack<B, D>(lds, ts, blockIdx.x, threadIdx.x);
    const uint32_t n = W * H;
    const uint32_t G = gridDim.x * B;
    const uint32_t words = s.hit.compressed ? 2u : 4u;
    for (uint32_t i = blockIdx.x * B + threadIdx.x; i < n; i += G) {
        f3 o, d;
        float tmin, tmax;
        primaryRay(s.cam, i % W, i / W, W, H, o, d, tmin, tmax, jx, jy);
        HitRec h;
        const bool hit = traceBVH4<false, false>(s, o, d, tmin, tmax, st, h);
        uint32_t w[4] = {0u, 0u, 0u, 0u};
        float dep = 1.f, msk = 0.f;
        if (hit) {
            const uint32_t m = meshOfTriangle(s, h.tri);
            packTriangleHit(s.hit, m, h.tri - s.meshTriOff[m], h.b1, h.b2, w);
            const uint32_t* tv = s.idx + 3 * size_t(h.tri);
            const float4 p0 = s.vtx[2 * size_t(tv[0])], p1 = s.vtx[2 * size_t(tv[1])], p2 = s.vtx[2 * size_t(tv[2])];
            const float b0 = 1.f - h.b1 - h.b2;
            const f3 P = mk3(p0.x, p0.y, p0.z) * b0 + mk3(p1.x, p1.y, p1.z) * h.b1 + mk3(p2.x, p2.y, p2.z) * h.b2;
            const float4 z = s.viewProjZ, ww = s.viewProjW;
            dep = (z.x * P.x + z.y * P.y + z.z * P.z + z.w) / (ww.x * P.x + ww.y * P.y + ww.z * P.z + ww.w);
            msk = 1.f;
        }
        if (words == 2u) {
            reinterpret_cast<uint2*>(vbuf)[i] = make_uint2(w[0], w[1]);
        } else {
            reinterpret_cast<uint4*>(vbuf)[i] = make_uint4(w[0], w[1], w[2], w[3]);
        }
        if (depth) depth[i] = dep;
        if (viewW) viewW[i] = make_float4(-d.x, -d.y, -d.z, 0.f);
        if (mask) mask[i] = msk;
    }
}

hipError_t launchVBuffer(const SceneDev& s, uint32_t W, uint32_t H, float jx, float jy, uint32_t* vbuf, float* depth,
                         float4* viewW, float* mask, const TraceScratch& ts, hipStream_t stream) {
    const uint32_t n = W * H;
    if (n == 0) return hipSuccess;
    return withTraceConfig(ts, [&](auto Bc, auto Dc) {
        constexpr int B = decltype(Bc)::value, D = decltype(Dc)::value;
        const uint32_t blocks = std::min<uint32_t>(ts.gridBlocks, (n + B - 1) / B);
        hipLaunchKernelGGL((kVBuffer<B, D>), dim3(blocks), dim3(B), 0, stream, s, W, H, jx, jy, vbuf, depth, viewW, mask, ts);
        return hipGetLastError();
    });
}

// Primary hits of a bound V-buffer: PackedHitInfo -> (triangle, 0, b1, b2) per pixel.
__global__ void kVBufferHits(SceneDev s, MptArgs a) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.nPix) return;
    const uint32_t words = s.hit.compressed ? 2u : 4u;
    uint32_t w[4] = {0u, 0u, 0u, 0u};
    if (words == 2u) {
        const uint2 v = reinterpret_cast<const uint2*>(a.vbuf)[p];
        w[0] = v.x;
        w[1] = v.y;
    } else {
        const uint4 v = reinterpret_cast<const uint4*>(a.vbuf)[p];
        w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
    }
    uint32_t tri;
    float b1, b2;
    a.pixHit[p] = unpackTriangleHit(s, w, tri, b1, b2) ? make_uint4(tri, 0u, f2u(b1), f2u(b2)) : make_uint4(kInvalid, 0u, 0u, 0u);
}

template <bool MULTI>
__global__ void kPrimaryRays(SceneDev s, MptArgs a) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.nPix) return;
    const ViewCtx c = viewCtx<MULTI>(s, a, p, false);
    const uint32_t q = p - c.pixBase;
    f3 o, d;
    float tmin, tmax;
    primaryRay(*c.cam, q % c.W, q / c.W, c.W, c.H, o, d, tmin, tmax);
    a.rayQ[0][2 * p] = make_float4(o.x, o.y, o.z, tmin);
    a.rayQ[0][2 * p + 1] = make_float4(d.x, d.y, d.z, tmax);
}

template <int B, int D, bool ANY, bool COUNT>
__global__ __launch_bounds__(B) void kTraceQueue(SceneDev s, const float4* __restrict__ rays, const uint32_t* countPtr,
                                                 uint32_t countConst, uint4* __restrict__ hits, TraceScratch ts,
                                                 unsigned long long* stats) {
    __shared__ uint32_t lds[D * B];
    TraceStack st = makeStack<B, D>(lds, ts, blockIdx.x, threadIdx.x);
    const uint32_t n = countPtr ? *countPtr : countConst;
    const uint32_t G = gridDim.x * B;
    TraceCounters cnt{0, 0};
    for (uint32_t i = blockIdx.x * B + threadIdx.x; i < n; i += G) {
        const float4 ra = rays[2 * i], rb = rays[2 * i + 1];
        HitRec h;
        const bool hit = traceBVH4<ANY, COUNT>(s, mk3(ra.x, ra.y, ra.z), mk3(rb.x, rb.y, rb.z), ra.w, rb.w, st, h, &cnt);
        hits[i] = hit ? make_uint4(h.tri, f2u(h.t), f2u(h.b1), f2u(h.b2)) : make_uint4(kInvalid, 0u, 0u, 0u);
    }
    if (COUNT) {
        atomicAdd(&stats[0], (unsigned long long)cnt.nodes);
        atomicAdd(&stats[1], (unsigned long long)cnt.tris);
    }
}

// Emissive next-event estimation at a shading point (extension of MinimalPathTracer; semantics of
// PTUtils.slang:177-249 without reparameterization). Produces an unoccluded contribution and the
// shadow ray that gates it.
// `sel` is the probability of having chosen emissive lights among the light types.
__device__ bool neeSample(const SceneDev& s, uint32_t kind, const Shading& sd, const Material& m, Rng& rng, f3& org, f3& dir,
                          float& dist, f3& contrib, float sel) {
    LightSample ls;
    if (!sampleLight(s, int(kind), sd.posW, sd.N, rng, ls)) return false;
    if (!(ls.pdf > 0.f)) return false;
    f3 Li = ls.Le / ls.pdf;
    if (!anyPos(Li)) return false;
    ls.pdf *= sel;
    Li = Li / sel;
    org = rayOriginOffset(sd.posW, dot(sd.faceN, ls.dir) >= 0.f ? sd.faceN : -sd.faceN);
    f3 lightPos = rayOriginOffset(ls.posW, ls.normalW);
    f3 toL = lightPos - org;
    dist = length(toL);
    dir = normalize(toL);
    f3 f = matEval(sd, m, dir);
    if (!anyPos(f)) return false;
    float mis = misPower(ls.pdf, matPdf(sd, m, dir));
    contrib = f * Li * mis;
    return true;
}

// Env map next-event estimation (extension of MinimalPathTracer; PathTracer.slang generateEnvMapSample
// with the NEE contribution and MIS of PathTracer.slang:695-735): one EnvMapSampler sample, a
// visibility ray to infinity from the side of the surface the direction leaves.
__device__ bool envNeeSample(const SceneDev& s, const Shading& sd, const Material& m, Rng& rng, f3& org, f3& dir, float& dist,
                             f3& contrib, float sel) {
    const float ux = rng.u1(), uy = rng.u1();
    f3 Le;
    float pdf;
    if (!envSample(s, ux, uy, dir, pdf, Le)) return false;
    f3 Li = pdf > 0.f ? Le / pdf : mk3(0.f);
    if (!anyPos(Li)) return false;
    pdf *= sel;
    Li = Li / sel;
    org = rayOriginOffset(sd.posW, dot(sd.faceN, dir) >= 0.f ? sd.faceN : -sd.faceN);
    dist = kFltMax;
    const f3 f = matEval(sd, m, dir);
    if (!anyPos(f)) return false;
    contrib = f * Li * misPower(pdf, matPdf(sd, m, dir));
    return true;
}

// Light types of next-event estimation at a vertex: env map (useEnvNEE and a sampleable env
// light), emissive triangles (useEmissiveNEE), analytic lights (any in the scene), each chosen with
// equal probability (PathTracer.slang:643-663).
struct LightTypes {
    float pEnv, pEmissive, pAnalytic;
    uint32_t count;
};
__device__ __forceinline__ LightTypes lightTypes(const SceneDev& s, const MptArgs& a) {
    const bool env = a.envNee && s.envImp && (s.envFlags & kEnvLight);
    const bool em = a.nee != 0u, an = s.nALights != 0u;
    LightTypes t;
    t.count = uint32_t(env) + uint32_t(em) + uint32_t(an);
    const float inv = t.count ? 1.f / float(t.count) : 0.f;
    t.pEnv = env ? inv : 0.f;
    t.pEmissive = em ? inv : 0.f;
    t.pAnalytic = an ? inv : 0.f;
    return t;
}

// Light configurations the shading kernels are specialised for (MptArgs.lightMode): no light
// sampling, emissive NEE as the only type (the code of a single type, without selection), any mix.
constexpr int kLightsNone = 0, kLightsEmissive = 1, kLightsAny = 2;

// Direct light at a vertex with one visibility ray: selectLightType (PathTracer.slang:671-693),
// then the chosen type's sample. A number is drawn for the selection only when more than one
// type is enabled, so one type alone keeps its own sample sequence (analytic lights alone are
// MinimalPathTracer's evalDirectAnalytic, MinimalPathTracer.rt.slang:253-286).
template <int LM>
__device__ bool directSample(const SceneDev& s, const MptArgs& a, const Shading& sd, const Material& m, Rng& rng, f3& org,
                             f3& dir, float& dist, f3& contrib) {
    if (LM == kLightsNone) return false;
    if (LM == kLightsEmissive) return neeSample(s, a.lightKind, sd, m, rng, org, dir, dist, contrib, 1.f);
    const LightTypes t = lightTypes(s, a);
    if (t.count == 0u) return false;
    uint32_t type;
    float sel = 1.f;
    if (t.count == 1u) {
        type = t.pEnv > 0.f ? 0u : (t.pEmissive > 0.f ? 1u : 2u);
    } else {
        float u = rng.u1();
        if (u < t.pEnv) { type = 0u; sel = t.pEnv; }
        else if ((u -= t.pEnv) < t.pEmissive) { type = 1u; sel = t.pEmissive; }
        else if ((u -= t.pEmissive) < t.pAnalytic) { type = 2u; sel = t.pAnalytic; }
        else return false;
    }
    if (type == 0u) return envNeeSample(s, sd, m, rng, org, dir, dist, contrib, sel);
    if (type == 1u) return neeSample(s, a.lightKind, sd, m, rng, org, dir, dist, contrib, sel);
    if (!analyticSample(s, sd, m, rng.u1(), org, dir, dist, contrib)) return false;
    contrib = contrib / sel;
    return true;
}

__device__ __forceinline__ void pushShadow(const MptArgs& a, uint32_t slot, f3 o, f3 d, float tmax, f3 c, uint32_t pathId,
                                           uint32_t flags) {
    float4* q = a.shQ + 3 * size_t(slot);
    q[0] = make_float4(o.x, o.y, o.z, u2f(pathId));
    q[1] = make_float4(d.x, d.y, d.z, tmax);
    q[2] = make_float4(c.x, c.y, c.z, u2f(flags));
}
__device__ __forceinline__ void pushRay(const MptArgs& a, int q, uint32_t slot, f3 o, f3 d, uint32_t pathId) {
    a.rayQ[q][2 * size_t(slot)] = make_float4(o.x, o.y, o.z, 0.f);
    a.rayQ[q][2 * size_t(slot) + 1] = make_float4(d.x, d.y, d.z, kFltMax);
    a.pathQ[q][slot] = pathId;
}

// Paths whose pixel has a primary hit, appended to pathQ[1] (free until the first bounce's
// shading) with their count at counters[kPrimaryHitCount]: one thread per pixel appends its spp
// path ids, so the primary shading runs on paths with a hit only and the resolve writes the
// background of the other pixels. Launched with kCompactBlock threads per block (large blocks:
// one atomic per block).
constexpr int kShadeBlock = 256, kCompactBlock = 1024;
template <bool MULTI>
__global__ __launch_bounds__(kCompactBlock) void kCompactPrimary(SceneDev s0, MptArgs a) {
    __shared__ uint32_t lds[kCompactBlock / 64 + 1];
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t q = idx;
    bool inFrame = idx < a.nPix;
    if (!MULTI && a.W > 0u) {
        // A single view's pixels in 8x8 tiles (the primary traversal's order), so the primary
        // shading and the rays it starts work on compact patches of the image.
        const uint32_t tilesX = (a.W + 7u) / 8u, t = idx >> 6;
        const uint32_t x = (t % tilesX) * 8u + (idx & 7u), y = (t / tilesX) * 8u + ((idx >> 3) & 7u);
        inFrame = x < a.W && y < a.H;
        q = y * a.W + x;
    }
    uint32_t cnt = 0, first = 0, view = 0;
    if (inFrame && a.pixHit[q].x != kInvalid) {
        if (MULTI) view = findView(a.views, a.nViews, q, false);
        const ViewCtx c = viewCtx<MULTI>(s0, a, q, false);
        cnt = c.spp;
        first = c.pathBase + (q - c.pixBase) * c.spp;
    }
    const uint32_t slot = blockAppendN<kCompactBlock>(cnt, &a.counters[kPrimaryHitCount], lds);
    for (uint32_t k = 0; k < cnt; k++) {
        a.pathQ[1][slot + k] = first + k;
        if (MULTI) a.pathView[first + k] = uint16_t(view);  // read by every later per-path view lookup
    }
}

// The shading kernels are held to 128 VGPRs (4 waves per SIMD instead of 3 at their natural
// 134-142): the light-BVH descent is a chain of dependent loads, and the extra wave hides more of
// it than the few spilled values cost (the emissive-only instances spill nothing).
//
// Primary-hit shading (MinimalPathTracer.rt.slang:392-449), one thread per path with a primary
// hit (the compacted list of kCompactPrimary). The path state (throughput, last normal, sample
// generator) is written for paths that continue only: nothing reads it for the others.
template <bool MULTI, int LM>
__global__ __launch_bounds__(kShadeBlock) __attribute__((amdgpu_waves_per_eu(4))) void kShadePrimary(SceneDev s0, MptArgs a) {
    __shared__ uint32_t lds[kShadeBlock / 64 + 1];
    const uint32_t n = a.counters[kPrimaryHitCount];
    const uint32_t G = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += G) {
        const uint32_t i = base + threadIdx.x;
        bool cont = false, shadow = false;
        f3 newO, newD, shO, shD, shC;
        float shT = 0.f;
        uint32_t p = 0;
        if (i < n) {
            p = a.pathQ[1][i];
            const ViewCtx c = viewCtx<MULTI>(s0, a, p, true);
            const SceneDev& s = *c.s;
            const uint32_t lp = p - c.pathBase;
            const uint32_t pix = lp / c.spp, k = lp % c.spp;
            const uint32_t px = pix % c.W, py = pix / c.W;
            const uint4 h = a.pixHit[c.pixBase + pix];
            f3 o, d;
            float tmin, tmax;
            if (!MULTI && a.viewW) {
                // getPrimaryRayDir with a bound viewW (MinimalPathTracer.rt.slang: -viewW).
                const float4 v = a.viewW[pix];
                d = mk3(-v.x, -v.y, -v.z);
            } else {
                primaryRay(*c.cam, px, py, c.W, c.H, o, d, tmin, tmax);
            }
            Shading sd = shadingFromHit(s, h.x, u2f(h.z), u2f(h.w), d);
            Material m = loadMaterial(s, s.mats, s.texels, sd, sd, nullptr);
            Rng rng = rngSeed(px, py, c.frame * c.spp + k);
            f3 org = rayOriginFromShading(sd);
            f3 Lp = mk3(0.f);
            if (a.computeDirect) {
                Lp = Lp + m.emission;
                shadow = directSample<LM>(s, a, sd, m, rng, shO, shD, shT, shC);
            }
            Scatter sc;
            if (matSample(sd, m, rng, sc, a.importance != 0)) {
                const f3 thp = mk3(1.f) * sc.weight;
                cont = anyPos(thp);
                newO = sc.transmission ? rayOriginFromShading(sd, false) : org;  // computeRayOrigin(false) for transmission
                newD = sc.wo;
                if (cont) {
                    a.thp[p] = make_float4(thp.x, thp.y, thp.z, sc.pdf);
                    a.prevN[p] = make_float4(sd.N.x, sd.N.y, sd.N.z, u2f(sc.delta ? 1u : 0u));
                    a.rng[p] = make_uint4(rng.s0, rng.s1, rng.s2, rng.s3);
                }
            }
            a.Lp[p] = make_float4(Lp.x, Lp.y, Lp.z, 0.f);
            a.Ls[p] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        const uint32_t slot = blockAppend<kShadeBlock>(cont, &a.counters[0], lds);
        if (cont) pushRay(a, 0, slot, newO, newD, p);
        const uint32_t sslot = blockAppend<kShadeBlock>(shadow, &a.counters[kCounterStride + 0], lds);
        if (shadow) pushShadow(a, sslot, shO, shD, shT, shC, p, 1u);
    }
}

// True when the hits of a depth (n of them) are few enough for the tail kernel, which then runs
// their paths to the end; the per-depth shading launch leaves them alone.
__device__ __forceinline__ bool tailActive(const MptArgs& a, uint32_t n) { return a.tailPaths != 0u && n <= a.tailPaths; }

// handleHit for scatter depth `depth` (MinimalPathTracer.rt.slang:324-376), one thread per queued ray.
// Runs on the compacted hits of the bounce (entry k: hit hitQ[k] of ray hitIdx[k]) unless
// visibility rays have launches of their own, then on every queued ray (hitQ[i] of ray i).
template <bool MULTI, int LM>
__global__ __launch_bounds__(kShadeBlock) __attribute__((amdgpu_waves_per_eu(4))) void kShade(SceneDev s0, MptArgs a, uint32_t depth) {
    __shared__ uint32_t lds[kShadeBlock / 64 + 1];
    const int qi = depth & 1, qo = qi ^ 1;
    const bool compact = !a.separateShadow;
    const uint32_t n = compact ? a.counters[kHitCount + depth] : a.counters[depth];
    if (compact && tailActive(a, n)) return;
    const uint32_t G = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += G) {
        const uint32_t k = base + threadIdx.x;
        bool cont = false, shadow = false;
        f3 newO, newD, shO, shD, shC;
        float shT = 0.f;
        uint32_t pathId = 0;
        if (k < n) {
            const uint32_t i = compact ? a.hitIdx[k] : k;
            pathId = a.pathQ[qi][i];
            const uint4 h = a.hitQ[k];
            const SceneDev& s = *viewCtx<MULTI>(s0, a, pathId, true).s;
            if (h.x != kInvalid) {
                const float4 ro = a.rayQ[qi][2 * size_t(i)], rd = a.rayQ[qi][2 * size_t(i) + 1];
                const f3 o = mk3(ro.x, ro.y, ro.z), d = mk3(rd.x, rd.y, rd.z);
                float4 t4 = a.thp[pathId], n4 = a.prevN[pathId];
                f3 thp = mk3(t4.x, t4.y, t4.z);
                const bool prevDelta = (f2u(n4.w) & 1u) != 0u;
                uint4 r4 = a.rng[pathId];
                Rng rng{r4.x, r4.y, r4.z, r4.w};
                Shading sd = shadingFromHit(s, h.x, u2f(h.z), u2f(h.w), d);
                Material m = loadMaterial(s, s.mats, s.texels, sd, sd, nullptr);
                float4 ls4 = a.Ls[pathId];
                f3 Ls = mk3(ls4.x, ls4.y, ls4.z);
                if (a.computeDirect || depth > 0) {
                    f3 e = m.emission;
                    if (a.nee && anyPos(e) && !prevDelta) {
                        float lp = lightPdf(s, int(a.lightKind), o, mk3(n4.x, n4.y, n4.z), s.triToLight[h.x], sd.posW,
                                            sd.orientedFaceN());
                        e = e * misPower(t4.w, LM == kLightsAny ? lp * lightTypes(s, a).pEmissive : lp);
                    }
                    Ls = Ls + thp * e;
                    a.Ls[pathId] = make_float4(Ls.x, Ls.y, Ls.z, 0.f);
                }
                if (depth < a.maxBounces) {
                    f3 org = rayOriginFromShading(sd);
                    shadow = directSample<LM>(s, a, sd, m, rng, shO, shD, shT, shC);
                    if (shadow) shC = thp * shC;
                    Scatter sc;
                    if (matSample(sd, m, rng, sc, a.importance != 0)) {
                        thp = thp * sc.weight;
                        cont = anyPos(thp);
                        newO = sc.transmission ? rayOriginFromShading(sd, false) : org;  // computeRayOrigin(false) for transmission
                        newD = sc.wo;
                        a.thp[pathId] = make_float4(thp.x, thp.y, thp.z, sc.pdf);
                        a.prevN[pathId] = make_float4(sd.N.x, sd.N.y, sd.N.z, u2f(sc.delta ? 1u : 0u));
                    }
                    a.rng[pathId] = make_uint4(rng.s0, rng.s1, rng.s2, rng.s3);
                }
            } else if (s.envTex && (s.envFlags & kEnvLight) && (a.computeDirect || depth > 0)) {
                // scatterMiss (MinimalPathTracer.rt.slang:475-486): the env map lights the missed ray.
                const float4 rd = a.rayQ[qi][2 * size_t(i) + 1];
                const float4 t4 = a.thp[pathId], ls4 = a.Ls[pathId];
                const f3 rdir = mk3(rd.x, rd.y, rd.z);
                f3 Le = envEval(s, rdir);
                // MIS against env map sampling at the previous vertex (PathTracer.slang:1096-1108).
                const float pEnv = LM == kLightsAny ? lightTypes(s, a).pEnv : 0.f;
                if (pEnv > 0.f && !(f2u(a.prevN[pathId].w) & 1u)) Le = Le * misPower(t4.w, envPdf(s, rdir) * pEnv);
                a.Ls[pathId] = make_float4(ls4.x + t4.x * Le.x, ls4.y + t4.y * Le.y, ls4.z + t4.z * Le.z, 0.f);
            }
        }
        const uint32_t slot = blockAppend<kShadeBlock>(cont, &a.counters[depth + 1], lds);
        if (cont) pushRay(a, qo, slot, newO, newD, pathId);
        const uint32_t sslot = blockAppend<kShadeBlock>(shadow, &a.counters[kCounterStride + depth + 1], lds);
        if (shadow) pushShadow(a, sslot, shO, shD, shT, shC, pathId, 0u);
    }
}

// Tail of the bounce loop. Late bounces hold few paths, and a launch per bounce then lasts as long
// as its slowest ray's chain of dependent node loads. Once a depth has at most tailPaths hits,
// this kernel takes them over (kShade of that depth returns at once) and runs each path to its end
// in one pair of lanes: handleHit, the visibility ray, the next closest-hit ray, and so on. The
// steps are kShade's and kTraceBounce's with the path state in registers, in the same order (the
// visibility ray of depth d adds its contribution after the emission of depth d and before that
// of depth d + 1; the library is built without implicit FMA contraction), so the radiance is
// bit-identical to the per-depth launches. Later depths find their queues empty. Rays traced here
// are counted in counters[kTailClosestCount] and counters[kTailVisibilityCount].
template <int B, int D, bool MULTI, int LM>
__global__ __launch_bounds__(B) void kTail(SceneDev s0, MptArgs a, uint32_t depth, TraceScratch ts) {
    __shared__ uint32_t lds[D * B];
    const uint32_t n = a.counters[kHitCount + depth];
    if (!tailActive(a, n)) return;
    TraceStack st = makeStack<B, D>(lds, ts, blockIdx.x, threadIdx.x);
    const int qi = depth & 1;
    // a.tailLanes paths per wave (at most 32; the other lanes idle): a wave's bounce lasts as
    // long as its slowest path's, and fewer paths per wave make that maximum smaller and let more
    // waves hide each other's latency. Each path has two lanes that shade it alike; then the even
    // lane traces the vertex's next closest-hit ray while the odd lane traces its visibility ray,
    // and each reads the other's answer, so a bounce waits for the longer of the two traversals
    // instead of both in turn.
    const uint32_t P = a.tailLanes - 1u < 32u ? a.tailLanes : 32u;
    const uint32_t lane = threadIdx.x & 63u, waves = (gridDim.x * B) >> 6;
    const uint32_t wave = (blockIdx.x * B + threadIdx.x) >> 6;
    const uint32_t role = lane & 1u, slot = lane >> 1;  // role 0: closest-hit ray, 1: visibility ray
    uint32_t nClosest = 0, nVis = 0;
    for (uint32_t k = slot < P ? wave * P + slot : n; k < n; k += waves * P) {
        const uint32_t i = a.hitIdx[k];
        const uint32_t pathId = a.pathQ[qi][i];
        const SceneDev& s = *viewCtx<MULTI>(s0, a, pathId, true).s;
        const GeomRef g{s.bvh, s.triPos};
        uint4 h = a.hitQ[k];
        const float4 ro = a.rayQ[qi][2 * size_t(i)], rd = a.rayQ[qi][2 * size_t(i) + 1];
        f3 o = mk3(ro.x, ro.y, ro.z), d = mk3(rd.x, rd.y, rd.z);
        float4 t4 = a.thp[pathId], n4 = a.prevN[pathId];
        const float4 ls4 = a.Ls[pathId];
        f3 Ls = mk3(ls4.x, ls4.y, ls4.z);
        const uint4 r4 = a.rng[pathId];
        Rng rng{r4.x, r4.y, r4.z, r4.w};
        for (uint32_t dep = depth;; dep++) {
            f3 thp = mk3(t4.x, t4.y, t4.z);
            if (h.x == kInvalid) {
                // scatterMiss (MinimalPathTracer.rt.slang:475-486) with MIS against env sampling.
                if (s.envTex && (s.envFlags & kEnvLight) && (a.computeDirect || dep > 0)) {
                    f3 Le = envEval(s, d);
                    const float pEnv = LM == kLightsAny ? lightTypes(s, a).pEnv : 0.f;
                    if (pEnv > 0.f && !(f2u(n4.w) & 1u)) Le = Le * misPower(t4.w, envPdf(s, d) * pEnv);
                    Ls = mk3(Ls.x + t4.x * Le.x, Ls.y + t4.y * Le.y, Ls.z + t4.z * Le.z);
                }
                break;
            }
            Shading sd = shadingFromHit(s, h.x, u2f(h.z), u2f(h.w), d);
            Material m = loadMaterial(s, s.mats, s.texels, sd, sd, nullptr);
            if (a.computeDirect || dep > 0) {
                f3 e = m.emission;
                if (a.nee && anyPos(e) && !(f2u(n4.w) & 1u)) {
                    float lp = lightPdf(s, int(a.lightKind), o, mk3(n4.x, n4.y, n4.z), s.triToLight[h.x], sd.posW,
                                        sd.orientedFaceN());
                    e = e * misPower(t4.w, LM == kLightsAny ? lp * lightTypes(s, a).pEmissive : lp);
                }
                Ls = Ls + thp * e;
            }
            if (dep >= a.maxBounces) break;
            const f3 org = rayOriginFromShading(sd);
            f3 shO, shD, shC;
            float shT = 0.f;
            const bool shadow = directSample<LM>(s, a, sd, m, rng, shO, shD, shT, shC);
            if (shadow) shC = thp * shC;
            bool cont = false;
            f3 newO, newD;
            Scatter sc;
            if (matSample(sd, m, rng, sc, a.importance != 0)) {
                thp = thp * sc.weight;
                cont = anyPos(thp);
                newO = sc.transmission ? rayOriginFromShading(sd, false) : org;
                newD = sc.wo;
                t4 = make_float4(thp.x, thp.y, thp.z, sc.pdf);
                n4 = make_float4(sd.N.x, sd.N.y, sd.N.z, u2f(sc.delta ? 1u : 0u));
            }
            // Both lanes of the path hold the same shadow, cont and rays: they trace one each.
            HitRec hr{kInvalid, 0.f, 0.f, 0.f};
            bool found = false;
            if (role ? shadow : cont)  // one call site: the pair's two traversals run side by side
                found = traceBVH4<false, false>(g, role ? shO : newO, role ? shD : newD, 0.f, role ? shT : kFltMax, st,
                                                hr, nullptr, role != 0u);
            const bool occluded = __shfl(found ? 1u : 0u, int(lane | 1u)) != 0u;
            const uint32_t htri = __shfl(hr.tri, int(lane & ~1u));
            const uint32_t ht = __shfl(f2u(hr.t), int(lane & ~1u)), hb1 = __shfl(f2u(hr.b1), int(lane & ~1u));
            const uint32_t hb2 = __shfl(f2u(hr.b2), int(lane & ~1u));
            if (shadow) {
                nVis += role ^ 1u;
                if (!occluded) Ls = Ls + shC;
            }
            if (!cont) break;
            nClosest += role ^ 1u;
            if (htri == kInvalid && !a.keepMisses) break;
            h = make_uint4(htri, ht, hb1, hb2);
            o = newO;
            d = newD;
        }
        if (role == 0u) a.Ls[pathId] = make_float4(Ls.x, Ls.y, Ls.z, 0.f);
    }
    for (int off = 32; off > 0; off >>= 1) {
        nClosest += __shfl_xor(nClosest, off);
        nVis += __shfl_xor(nVis, off);
    }
    if ((threadIdx.x & 63u) == 0u) {
        if (nClosest) atomicAdd(&a.counters[kTailClosestCount], nClosest);
        if (nVis) atomicAdd(&a.counters[kTailVisibilityCount], nVis);
    }
}

// Shadow rays: any hit in [0, dist]; an unoccluded ray adds its contribution to its path.
template <int B, int D, bool COUNT>
__global__ __launch_bounds__(B) void kShadow(SceneDev s, MptArgs a, uint32_t slotIdx, TraceScratch ts,
                                             unsigned long long* stats) {
    __shared__ uint32_t lds[D * B];
    TraceStack st = makeStack<B, D>(lds, ts, blockIdx.x, threadIdx.x);
    const uint32_t n = a.counters[kCounterStride + slotIdx];
    const uint32_t G = gridDim.x * B;
    TraceCounters cnt{0, 0};
    for (uint32_t i = blockIdx.x * B + threadIdx.x; i < n; i += G) {
        const float4* q = a.shQ + 3 * size_t(i);
        const float4 q0 = q[0], q1 = q[1];
        HitRec h;
        bool occ = traceBVH4<true, COUNT>(s, mk3(q0.x, q0.y, q0.z), mk3(q1.x, q1.y, q1.z), 0.f, q1.w, st, h, &cnt);
        if (!occ) {
            const float4 q2 = q[2];
            const uint32_t pathId = f2u(q0.w);
            float4* L = (f2u(q2.w) & 1u) ? &a.Lp[pathId] : &a.Ls[pathId];
            float4 v = *L;
            *L = make_float4(v.x + q2.x, v.y + q2.y, v.z + q2.z, 0.f);
        }
    }
    if (COUNT) {
        atomicAdd(&stats[0], (unsigned long long)cnt.nodes);
        atomicAdd(&stats[1], (unsigned long long)cnt.tris);
    }
}

// One traversal launch per bounce: the visibility rays queued for a bounce (count at nSPtr, none
// when null) and its closest-hit rays (count at nCPtr, or nCConst when null) share a persistent
// grid. Visibility rays come first in the index space; a wave mixes the two kinds only at the
// boundary. Closest hits go to hits[j], unoccluded visibility rays add their contribution to
// their path.
//
// Closest hits go to hits[j] for every ray j when hitCount is null (primary rays, one per
// pixel); otherwise only rays that hit something are appended (hits[k], hitIdx[k] = j, count at
// *hitCount), so the shading launch runs on dense lanes. A miss needs no shading: the scenes have
// no environment light (MinimalPathTracer.rt.slang:378-390 adds nothing for a black background).
//
// In a batch (MULTI) each ray traverses the scene of its view: visibility and bounce rays find it
// through their path id (pathIds[j] for bounce ray j), primary rays through their pixel.
//
// The register budget is held to 8 waves per SIMD (64 VGPRs): with a 16-entry LDS stack a CU
// then holds 8 blocks of 256 threads, and the extra waves hide more of the node fetch latency
// than the registers given up cost (a 24-entry stack is limited to 6 blocks by LDS either way).
template <int B, int D, bool COUNT, bool MULTI>
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(8))) void kTraceBounce(
    SceneDev s, MptArgs a, const float4* __restrict__ rays, const uint32_t* nCPtr, uint32_t nCConst,
    uint4* __restrict__ hits, uint32_t* __restrict__ hitIdx, uint32_t* hitCount, const uint32_t* nSPtr,
    const uint32_t* __restrict__ pathIds, TraceScratch ts, unsigned long long* stats) {
    __shared__ uint32_t lds[D * B];
    TraceStack st = makeStack<B, D>(lds, ts, blockIdx.x, threadIdx.x);
    const uint32_t nS = nSPtr ? *nSPtr : 0u;
    const uint32_t n = nS + (nCPtr ? *nCPtr : nCConst);
    const uint32_t G = gridDim.x * B;
    TraceCounters cc{0, 0}, cs{0, 0};
    // R consecutive rays per wave: all 64 lanes while the rays fill the grid; fewer rays than
    // lanes are spread over more waves (more waves per SIMD to hide latency, and a wave waits for
    // the slowest of fewer rays) instead of filling the first waves, but at least 32 per wave:
    // thinner waves compete for the same VALU issue slots with most of their lanes idle.
    // A single view's primary rays (one per pixel, their hits stored by pixel) are visited in 8x8
    // pixel tiles: a wave's rays then leave a compact patch of the image instead of a 64-pixel
    // strip of one row. Index i runs over whole tiles; positions outside the frame are skipped.
    const bool tiled = !MULTI && !hitCount && !nSPtr && !pathIds && a.W > 0u;
    const uint32_t tilesX = (a.W + 7u) / 8u;
    const uint32_t nI = tiled ? tilesX * ((a.H + 7u) / 8u) * 64u : n;
    const uint32_t waves = G >> 6, lane = threadIdx.x & 63u, wave = (blockIdx.x * B + threadIdx.x) >> 6;
    const uint32_t R = nI >= G ? 64u : max((nI + waves - 1u) / waves, 32u);
    for (uint32_t i = lane < R ? wave * R + lane : nI; i < nI; i += waves * R) {
        uint32_t pix = i;
        if (tiled) {
            const uint32_t t = i >> 6, x = (t % tilesX) * 8u + (i & 7u), y = (t / tilesX) * 8u + ((i >> 3) & 7u);
            if (x >= a.W || y >= a.H) continue;
            pix = y * a.W + x;
        }
        // Ray i: visibility record i (i < nS) or closest-hit ray i - nS (primary rays: pixel pix).
        const bool vis = pix < nS;
        const uint32_t j = vis ? pix : pix - nS;
        const float4* q = vis ? a.shQ + 3 * size_t(j) : rays + 2 * size_t(j);
        const float4 r0 = q[0], r1 = q[1];
        GeomRef g{s.bvh, s.triPos};
        if (MULTI) {
            // Visibility and bounce rays belong to paths with a primary hit (view index stored per
            // path); primary rays (no path ids) are found by pixel.
            const bool byPath = vis || pathIds != nullptr;
            const uint32_t key = vis ? f2u(r0.w) : (pathIds ? pathIds[j] : j);
            const uint32_t view = byPath ? uint32_t(a.pathView[key]) : findView(a.views, a.nViews, key, false);
            const SceneDev& sv = a.scenes[a.views[view].scene];
            g = GeomRef{sv.bvh, sv.triPos};
        }
        TravState r;
        r.start(mk3(r0.x, r0.y, r0.z), mk3(r1.x, r1.y, r1.z), vis ? 0.f : r0.w, r1.w, vis);
        TraceCounters c{0, 0};
        while (r.step<COUNT>(g, st, &c)) {
        }
        if (COUNT) {
            TraceCounters& t = vis ? cs : cc;
            t.nodes += c.nodes;
            t.tris += c.tris;
        }
        const bool hit = r.hit.tri != kInvalid;
        if (vis && !hit) {
            const float4 r2 = q[2];
            const uint32_t pathId = f2u(r0.w);
            float4* L = (f2u(r2.w) & 1u) ? &a.Lp[pathId] : &a.Ls[pathId];
            const float4 v = *L;
            *L = make_float4(v.x + r2.x, v.y + r2.y, v.z + r2.z, 0.f);
        }
        const HitRec& h = r.hit;
        if (!hitCount) {
            if (!vis) hits[j] = hit ? make_uint4(h.tri, f2u(h.t), f2u(h.b1), f2u(h.b2)) : make_uint4(kInvalid, 0u, 0u, 0u);
        } else {
            const bool keep = !vis && (hit || a.keepMisses);
            const uint32_t k = waveAppend(keep, hitCount);
            if (keep) {
                hits[k] = make_uint4(h.tri, f2u(h.t), f2u(h.b1), f2u(h.b2));
                hitIdx[k] = j;
            }
        }
    }
    if (COUNT) {
        atomicAdd(&stats[0], (unsigned long long)cc.nodes);
        atomicAdd(&stats[1], (unsigned long long)cc.tris);
        atomicAdd(&stats[2], (unsigned long long)cs.nodes);
        atomicAdd(&stats[3], (unsigned long long)cs.tris);
    }
}

// outColor + radiance per sample, averaged over samples like AccumulatePass Single mode. A pixel
// without a primary hit shows the background (MinimalPathTracer.rt.slang:462-465: the env map
// seen along the primary ray, else black) in each of its samples.
template <bool MULTI>
__global__ void kResolve(SceneDev s0, MptArgs a, float4* out) {
    const uint32_t bpix = blockIdx.x * blockDim.x + threadIdx.x;
    if (bpix >= a.nPix) return;
    uint32_t pix = bpix, pathBase = 0, spp = a.spp, W = a.W, H = a.H;
    const SceneDev* s = &s0;
    const CameraDev* cam = &s0.cam;
    if (MULTI) {
        const ViewDev& v = a.views[findView(a.views, a.nViews, bpix, false)];
        pix = bpix - v.pixBase;
        pathBase = v.pathBase;
        spp = v.spp;
        out = v.out;
        W = v.W;
        H = v.H;
        s = a.scenes + v.scene;
        cam = &v.cam;
    }
    f3 sum = mk3(0.f);
    if (a.pixHit[bpix].x != kInvalid) {
        for (uint32_t k = 0; k < spp; k++) {
            const uint32_t p = pathBase + pix * spp + k;
            const float4 lp = a.Lp[p], ls = a.Ls[p];
            sum = sum + (mk3(lp.x, lp.y, lp.z) + mk3(ls.x, ls.y, ls.z));
        }
    } else {
        f3 bg = mk3(0.f);
        if (s->envTex && (s->envFlags & kEnvBackground)) {
            f3 o, d;
            float tmin, tmax;
            if (!MULTI && a.viewW) {
                const float4 v = a.viewW[pix];
                d = mk3(-v.x, -v.y, -v.z);
            } else {
                primaryRay(*cam, pix % W, pix / W, W, H, o, d, tmin, tmax);
            }
            bg = envEval(*s, d);
        }
        for (uint32_t k = 0; k < spp; k++) sum = sum + (bg + mk3(0.f));
    }
    const float w = 1.0f / float(spp);
    out[pix] = make_float4(sum.x * w, sum.y * w, sum.z * w, 1.f);
}

__global__ void kRayCounts(const MptArgs a, uint32_t nb, unsigned long long* out) {
    // out[0] += closest-hit rays (primary + all scatter queues), out[1] += shadow rays; out[6],
    // out[7] += the closest-hit and shadow rays of those that the tail launch traced.
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        unsigned long long c = a.vbuf ? 0ull : (unsigned long long)a.nPix, v = 0;
        for (uint32_t b = 0; b < nb; b++) c += a.counters[b];
        for (uint32_t b = 0; b <= nb; b++) v += a.counters[kCounterStride + b];
        c += a.counters[kTailClosestCount];
        v += a.counters[kTailVisibilityCount];
        out[0] += c;
        out[1] += v;
        out[6] += a.counters[kTailClosestCount];
        out[7] += a.counters[kTailVisibilityCount];
    }
}

__global__ void kSetU32(uint32_t* p, uint32_t v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

hipError_t launchMptPrologue(const MptArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(kSetU32, dim3(1), dim3(1), 0, st, a.counters + 2 * kCounterStride, a.frame);
    return hipGetLastError();
}

hipError_t launchMptEpilogue(const SceneDev& s, const MptArgs& a, float4* out, unsigned long long* d_counts, hipStream_t st) {
    const uint32_t npix = a.nPix;
    hipLaunchKernelGGL(kResolve<false>, dim3((npix + 255) / 256), dim3(256), 0, st, s, a, out);
    if (d_counts) hipLaunchKernelGGL(kRayCounts, dim3(1), dim3(64), 0, st, a, a.maxBounces + 1, d_counts);
    return hipGetLastError();
}

hipError_t launchMptFrame(const SceneDev& s, const MptArgs& a, const TraceScratch& ts, float4* out, bool count,
                          unsigned long long* d_counts, hipStream_t st, KernelTimer* tm) {
    hipError_t e = launchMptPrologue(a, st);
    if (e != hipSuccess) return e;
    e = launchMptBounces(s, a, ts, count, d_counts, st, tm);
    if (e != hipSuccess) return e;
    return launchMptEpilogue(s, a, out, d_counts, st);
}

template <bool MULTI>
hipError_t mptBounces(const SceneDev& s, const MptArgs& a0, const TraceScratch& ts, bool count, unsigned long long* d_counts,
                      hipStream_t st, KernelTimer* tm) {
    // Counting frames trace every ray in the per-depth launches (their node and triangle counters).
    MptArgs a = a0;
    if (count || a.separateShadow) a.tailPaths = 0u;
    KernelTimer dummy;
    if (!tm) tm = &dummy;
    const uint32_t npix = a.nPix;
    const int T = kShadeBlock;
    hipLaunchKernelGGL(kZeroU32, dim3(1), dim3(2 * kCounterStride), 0, st, a.counters, 2u * kCounterStride);
    hipLaunchKernelGGL(kZeroU32, dim3(1), dim3(kCounterStride), 0, st, a.counters + kHitCount, uint32_t(kCounterStride));
    const bool vbuf = !MULTI && a.vbuf;
    if (vbuf) hipLaunchKernelGGL(kVBufferHits, dim3((npix + T - 1) / T), dim3(T), 0, st, s, a);
    else hipLaunchKernelGGL(kPrimaryRays<MULTI>, dim3((npix + T - 1) / T), dim3(T), 0, st, s, a);
    unsigned long long* tstat = count ? d_counts + 2 : nullptr;   // closest-hit nodes, triangles
    unsigned long long* sstat = count ? d_counts + 4 : nullptr;   // shadow nodes, triangles
    const TraceScratch& tsc = ts;
    auto traceQ = [&](const float4* rays, const uint32_t* cptr, uint32_t cconst, uint4* hits, uint32_t maxBlocks) {
        tm->begin(KK_TRACE, st);
        hipError_t e = withTraceConfig(tsc, [&](auto Bc, auto Dc) {
            constexpr int B = decltype(Bc)::value, D = decltype(Dc)::value;
            const uint32_t blocks = std::min<uint32_t>(tsc.gridBlocks, maxBlocks);
            if (count) hipLaunchKernelGGL((kTraceQueue<B, D, false, true>), dim3(blocks), dim3(B), 0, st, s, rays, cptr, cconst, hits, tsc, tstat);
            else hipLaunchKernelGGL((kTraceQueue<B, D, false, false>), dim3(blocks), dim3(B), 0, st, s, rays, cptr, cconst, hits, tsc, tstat);
            return hipGetLastError();
        });
        tm->end(st);
        return e;
    };
    auto shadow = [&](uint32_t slot) {
        tm->begin(KK_SHADOW, st);
        hipError_t e = withTraceConfig(tsc, [&](auto Bc, auto Dc) {
            constexpr int B = decltype(Bc)::value, D = decltype(Dc)::value;
            if (count) hipLaunchKernelGGL((kShadow<B, D, true>), dim3(tsc.gridBlocks), dim3(B), 0, st, s, a, slot, tsc, sstat);
            else hipLaunchKernelGGL((kShadow<B, D, false>), dim3(tsc.gridBlocks), dim3(B), 0, st, s, a, slot, tsc, sstat);
            return hipGetLastError();
        });
        tm->end(st);
        return e;
    };
    auto bounce = [&](const float4* rays, const uint32_t* nC, uint32_t nCConst, uint4* hits, uint32_t* hitIdx,
                      uint32_t* hitCount, const uint32_t* nS, const uint32_t* pathIds, uint32_t maxBlocks) {
        tm->begin(KK_TRACE, st);
        hipError_t e = withTraceConfig(tsc, [&](auto Bc, auto Dc) {
            constexpr int B = decltype(Bc)::value, D = decltype(Dc)::value;
            const uint32_t blocks = std::min<uint32_t>(tsc.gridBlocks, maxBlocks);
            if (count)
                hipLaunchKernelGGL((kTraceBounce<B, D, true, MULTI>), dim3(blocks), dim3(B), 0, st, s, a, rays, nC, nCConst,
                                   hits, hitIdx, hitCount, nS, pathIds, tsc, tstat);
            else
                hipLaunchKernelGGL((kTraceBounce<B, D, false, MULTI>), dim3(blocks), dim3(B), 0, st, s, a, rays, nC, nCConst,
                                   hits, hitIdx, hitCount, nS, pathIds, tsc, tstat);
            return hipGetLastError();
        });
        tm->end(st);
        return e;
    };
    auto tail = [&](uint32_t depth) {
        tm->begin(KK_OTHER, st);
        hipError_t e = withTraceConfig(tsc, [&](auto Bc, auto Dc) {
            constexpr int B = decltype(Bc)::value, D = decltype(Dc)::value;
            // One wave per tailLanes paths, within the persistent grid (the overflow stack's size).
            const uint32_t lanes = a.tailLanes - 1u < 32u ? a.tailLanes : 32u;
            const uint64_t threads = 64ull * ((uint64_t(a.tailPaths) + lanes - 1) / lanes);
            const uint32_t blocks = uint32_t(std::max<uint64_t>(1u, std::min<uint64_t>(tsc.gridBlocks, (threads + B - 1) / B)));
            if (a.lightMode == kLightsNone)
                hipLaunchKernelGGL((kTail<B, D, MULTI, kLightsNone>), dim3(blocks), dim3(B), 0, st, s, a, depth, tsc);
            else if (a.lightMode == kLightsEmissive)
                hipLaunchKernelGGL((kTail<B, D, MULTI, kLightsEmissive>), dim3(blocks), dim3(B), 0, st, s, a, depth, tsc);
            else
                hipLaunchKernelGGL((kTail<B, D, MULTI, kLightsAny>), dim3(blocks), dim3(B), 0, st, s, a, depth, tsc);
            return hipGetLastError();
        });
        tm->end(st);
        return e;
    };
    // Primary rays: one per pixel (the reference's V-buffer), shared by the pixel's samples;
    // with a bound V-buffer its hits are decoded instead.
    hipError_t e = hipSuccess;
    if (!vbuf)
        e = a.separateShadow ? traceQ(a.rayQ[0], nullptr, npix, a.pixHit, (npix + tsc.block - 1) / tsc.block)
                             : bounce(a.rayQ[0], nullptr, npix, a.pixHit, nullptr, nullptr, nullptr, nullptr,
                                      (npix + tsc.block - 1) / tsc.block);
    if (e != hipSuccess) return e;
    tm->begin(KK_SHADE, st);
    const uint32_t shadeBlocks = std::min<uint32_t>((a.nPaths + T - 1) / T, uint32_t(numCUs()) * 32);
    const uint32_t nCompact = MULTI ? npix : ((a.W + 7u) / 8u) * ((a.H + 7u) / 8u) * 64u;  // whole 8x8 tiles
    hipLaunchKernelGGL(kCompactPrimary<MULTI>, dim3((nCompact + kCompactBlock - 1) / kCompactBlock), dim3(kCompactBlock), 0, st,
                       s, a);
    if (a.lightMode == kLightsNone)
        hipLaunchKernelGGL((kShadePrimary<MULTI, kLightsNone>), dim3(shadeBlocks), dim3(T), 0, st, s, a);
    else if (a.lightMode == kLightsEmissive)
        hipLaunchKernelGGL((kShadePrimary<MULTI, kLightsEmissive>), dim3(shadeBlocks), dim3(T), 0, st, s, a);
    else
        hipLaunchKernelGGL((kShadePrimary<MULTI, kLightsAny>), dim3(shadeBlocks), dim3(T), 0, st, s, a);
    tm->end(st);
    for (uint32_t b = 0; b <= a.maxBounces; b++) {
        // Closest-hit rays of bounce b with the visibility rays produced alongside them.
        if ((e = (a.separateShadow ? traceQ(a.rayQ[b & 1], a.counters + b, 0u, a.hitQ, 0xffffffffu)
                                   : bounce(a.rayQ[b & 1], a.counters + b, 0u, a.hitQ, a.hitIdx, a.counters + kHitCount + b,
                                            a.shadowRays ? a.counters + kCounterStride + b : nullptr, a.pathQ[b & 1], 0xffffffffu))) !=
             hipSuccess)
            return e;
        if (a.separateShadow && a.shadowRays && (e = shadow(b)) != hipSuccess) return e;
        tm->begin(KK_SHADE, st);
        if (a.lightMode == kLightsNone) hipLaunchKernelGGL((kShade<MULTI, kLightsNone>), dim3(shadeBlocks), dim3(T), 0, st, s, a, b);
        else if (a.lightMode == kLightsEmissive)
            hipLaunchKernelGGL((kShade<MULTI, kLightsEmissive>), dim3(shadeBlocks), dim3(T), 0, st, s, a, b);
        else hipLaunchKernelGGL((kShade<MULTI, kLightsAny>), dim3(shadeBlocks), dim3(T), 0, st, s, a, b);
        tm->end(st);
        if (a.tailPaths && (e = tail(b)) != hipSuccess) return e;
    }
    return hipGetLastError();
}

hipError_t launchMptBounces(const SceneDev& s, const MptArgs& a, const TraceScratch& ts, bool count,
                            unsigned long long* d_counts, hipStream_t st, KernelTimer* tm) {
    return mptBounces<false>(s, a, ts, count, d_counts, st, tm);
}

// A batch of views in one launch sequence. The kernels read the view and scene tables from device
// memory, so the bounce loop can be captured once per batch layout and replayed with new tables.
hipError_t launchMptBatchBounces(const MptArgs& a, const TraceScratch& ts, bool count, unsigned long long* d_counts,
                                 hipStream_t st, KernelTimer* tm) {
    const SceneDev none{};
    return mptBounces<true>(none, a, ts, count, d_counts, st, tm);
}
hipError_t launchMptBatchEpilogue(const MptArgs& a, unsigned long long* d_counts, hipStream_t st) {
    const SceneDev none{};
    hipLaunchKernelGGL(kResolve<true>, dim3((a.nPix + 255) / 256), dim3(256), 0, st, none, a, nullptr);
    if (d_counts) hipLaunchKernelGGL(kRayCounts, dim3(1), dim3(64), 0, st, a, a.maxBounces + 1, d_counts);
    return hipGetLastError();
}
hipError_t launchMptBatch(const MptArgs& a, const TraceScratch& ts, bool count, unsigned long long* d_counts, hipStream_t st,
                          KernelTimer* tm) {
    hipError_t e = launchMptBatchBounces(a, ts, count, d_counts, st, tm);
    if (e != hipSuccess) return e;
    return launchMptBatchEpilogue(a, d_counts, st);
}

// ------------------------------------------------------------------ small utilities
__global__ void kFill(float* p, uint64_t n, float v) {
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    uint64_t G = uint64_t(gridDim.x) * blockDim.x;
    for (; i < n; i += G) p[i] = v;
}
__global__ void kFillRGBA(float4* p, uint64_t n, float4 v) {
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    uint64_t G = uint64_t(gridDim.x) * blockDim.x;
    for (; i < n; i += G) p[i] = v;
}
hipError_t launchFillRGBA(float4* p, uint64_t n, float r, float g, float b, float a, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(kFillRGBA, dim3(uint32_t(blocks)), dim3(256), 0, st, p, n, make_float4(r, g, b, a));
    return hipGetLastError();
}
hipError_t launchFill(float* p, uint64_t n, float v, hipStream_t st) {
    if (n == 0) return hipSuccess;
    uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(kFill, dim3(uint32_t(blocks)), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}
// Column sum over hash rows (AggregateGradients.cs.slang:50-58).
__global__ void kAggregate(const float* tmp, uint32_t hashSize, uint32_t dim, float* out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= dim) return;
    float s = 0.f;
    for (uint32_t h = 0; h < hashSize; h++) s += tmp[size_t(h) * dim + i];
    out[i] = s;
}
hipError_t launchAggregate(const float* tmp, uint32_t hashSize, uint32_t dim, float* out, hipStream_t st) {
    if (dim == 0) return hipSuccess;
    hipLaunchKernelGGL(kAggregate, dim3((dim + 255) / 256), dim3(256), 0, st, tmp, hashSize, dim, out);
    return hipGetLastError();
}

}  // namespace fhp
