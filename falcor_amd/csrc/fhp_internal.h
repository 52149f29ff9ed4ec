// Internal interfaces between the C-ABI host code and the HIP kernels of libfhp.
#pragma once
#include "fhp_scene.h"
#include <type_traits>

namespace fhp {

// Traversal launch configurations. Block size (threads), LDS stack depth (entries per lane) and
// grid size (blocks per CU) are chosen per pass (autotuned); these are the compiled variants.
constexpr int kDefaultBlock = 128;
constexpr int kDefaultLds = 24;
constexpr int kDefaultBlocksPerCU = 8;
inline bool validTraceConfig(uint32_t block, uint32_t lds) {
    return (block == 64 || block == 128 || block == 256) && (lds == 16 || lds == 24 || lds == 32);
}
constexpr int kMaxBounceSlots = 30;
constexpr int kCounterStride = 32;
// Counter words: [0, 32) closest-hit queue lengths by depth, [32, 64) visibility queue lengths by
// slot, [64] frame number, [kHitCount, +32) compacted hit counts by depth.
constexpr int kHitCount = 3 * kCounterStride;
// Paths with a primary hit (the primary shading's compacted list); the last hit-count slot,
// beyond any bounce depth, so the frame's counter reset clears it.
constexpr int kPrimaryHitCount = kHitCount + kCounterStride - 1;
// Closest-hit and visibility rays traced by the bounce loop's tail kernel (slots beyond any
// bounce depth: maxBounces <= kMaxBounceSlots - 2).
constexpr int kTailClosestCount = kHitCount + kCounterStride - 3;
constexpr int kTailVisibilityCount = kHitCount + kCounterStride - 2;
constexpr int kCounterWords = 4 * kCounterStride;  // counters: [0, 32) ray queues by depth, [32, 64) shadow queues

// One view of a batched frame: its scene (index into the batch's scene table), camera, frame
// shape and the first pixel / path of the view in the batch's pixel and path index spaces.
struct ViewDev {
    CameraDev cam;
    uint32_t W, H, spp, frame;
    uint32_t scene;
    uint32_t pixBase, pathBase;
    float4* out;
};

// Per-path state and queues of the wavefront forward pass. Slots are indexed by path id
// (pixel * spp + sample, views concatenated in a batch); queues are compacted each bounce.
struct MptArgs {
    uint32_t W, H, spp, frame, maxBounces;
    uint32_t computeDirect, importance, nee, lightKind;
    uint32_t nPaths;
    uint32_t envNee;          // 1: env map next-event estimation (useEnvNEE)
    uint32_t shadowRays;      // 1: shading emits visibility rays (NEE of any light type)
    uint32_t lightMode;       // shading specialisation: 0 no light sampling, 1 emissive NEE only, 2 any mix
    uint32_t separateShadow;  // 1: visibility rays in launches of their own (else shared with closest-hit)
                              // 0 also compacts each bounce's hits: hitQ[k] / hitIdx[k], count in counters[kHitCount + depth]
    float4* thp;     // thp.xyz, pdf of the last scatter
    float4* prevN;   // shading normal of the last scatter, flags (bit 0: delta lobe)
    uint4* rng;      // xoshiro128** state
    float4* Lp;      // radiance added at the primary hit (outColor of tracePath)
    float4* Ls;      // radiance gathered along the scatter chain (rayData.radiance)
    float4* rayQ[2]; // 2 float4 per ray: (o, tmin) (d, tmax)
    uint32_t* pathQ[2];
    uint4* hitQ;     // (tri, t, b1, b2)
    uint32_t* hitIdx;  // ray-queue index of each compacted hit
    float4* shQ;     // 3 float4 per shadow ray: (o, pathId) (d, tmax) (contrib, flags)
    uint4* pixHit;   // primary hits per pixel
    uint32_t* counters;   // [0, 64) queue counters, [64] frame number (read by the kernels so a
                          // captured graph can be replayed for any frame)
    // Batch mode (several views, possibly of different scenes, in one launch sequence): views
    // and scenes are device tables; W, H, spp and frame above are unused.
    const ViewDev* views;
    const SceneDev* scenes;
    uint16_t* pathView;   // view index of each path with a primary hit (written by the primary compaction)
    uint32_t nViews;
    uint32_t nPix;        // pixels over all views (W * H for one view)
    // Bound inputs of a single-view frame (MinimalPathTracer.cpp:48-59): the primary hits as a
    // V-buffer of PackedHitInfo in the scene's format (replaces the primary traversal) and the
    // view directions (viewW = -ray direction). Null when not bound.
    const uint32_t* vbuf;
    const float4* viewW;
    uint32_t keepMisses;  // 1: closest-hit misses are queued for shading too (environment light)
    uint32_t tailPaths;   // a depth with at most this many hits runs its paths to the end in one
                          // launch (kTail); 0: every depth has its own launches
    uint32_t tailLanes;   // paths per wave of that launch (1..32, two lanes each; the other lanes idle)
};

// Bits of a scene's trace error word (TraceScratch::err), read back by fhp_scene_trace_errors.
constexpr uint32_t kTraceErrStackOverflow = 1u;

struct TraceScratch {
    uint32_t* ovf;     // overflow stack (ovfDepth words per thread of the persistent grid)
    uint32_t* err;     // trace error word of the scene (kTraceErr* bits), or null
    uint32_t ovfDepth;
    uint32_t gridBlocks;  // persistent grid size in blocks
    uint32_t block;       // threads per block (64, 128, 256)
    uint32_t lds;         // LDS stack entries per lane (16, 24, 32)
};

// Calls fn(integral_constant<block>, integral_constant<lds>) for the configuration in ts.
template <int V> using IC = std::integral_constant<int, V>;
template <typename Fn> hipError_t withTraceConfig(const TraceScratch& ts, Fn&& fn) {
    switch (ts.block * 100 + ts.lds) {
        case 6416: return fn(IC<64>{}, IC<16>{});
        case 6424: return fn(IC<64>{}, IC<24>{});
        case 6432: return fn(IC<64>{}, IC<32>{});
        case 12816: return fn(IC<128>{}, IC<16>{});
        case 12824: return fn(IC<128>{}, IC<24>{});
        case 12832: return fn(IC<128>{}, IC<32>{});
        case 25616: return fn(IC<256>{}, IC<16>{});
        case 25624: return fn(IC<256>{}, IC<24>{});
        case 25632: return fn(IC<256>{}, IC<32>{});
        default: return hipErrorInvalidValue;
    }
}

struct KernelStats {
    unsigned long long* d;  // [nodes, tris] when counting is enabled
};

// Optional per-launch timing with HIP events on the launch stream, grouped by kernel kind.
enum KernelKind { KK_TRACE = 0, KK_SHADOW = 1, KK_SHADE = 2, KK_WAR_PRIMAL = 3, KK_WAR_ADJOINT = 4, KK_OTHER = 5,
                  KK_COUNT = 6 };
struct KernelTimer {
    static constexpr int kMax = 96;
    hipEvent_t ev[2 * kMax];
    int kind[kMax];
    int n = 0;
    bool on = false, created = false;
    double ms[KK_COUNT] = {};
    unsigned long long launches[KK_COUNT] = {};
    void begin(int k, hipStream_t st) {
        if (!on || n >= kMax) return;
        (void)hipEventRecord(ev[2 * n], st);
        kind[n] = k;
    }
    void end(hipStream_t st) {
        if (!on || n >= kMax) return;
        (void)hipEventRecord(ev[2 * n + 1], st);
        n++;
    }
};

// Launchers (fhp_kernels.hip). All are asynchronous on `stream`.
// VBufferRT (VBufferRT.rt.slang, VBufferRT.slang:56-143): primary hits as PackedHitInfo (2 or 4
// words per pixel in the scene's format), and optionally NDC depth, viewW and the hit mask.
// (jx, jy): camera jitter of the frame (Camera.slang:68-78).
hipError_t launchVBuffer(const SceneDev& s, uint32_t W, uint32_t H, float jx, float jy, uint32_t* vbuf, float* depth,
                         float4* viewW, float* mask, const TraceScratch& ts, hipStream_t stream);
hipError_t launchTraceRays(const SceneDev& s, const float4* rays, uint32_t n, uint4* hits, uint32_t* occluded, bool any,
                           const TraceScratch& ts, hipStream_t stream);
// Full frame: frame-number store, bounce loop, resolve, optional ray counts.
hipError_t launchMptBatch(const MptArgs& a, const TraceScratch& ts, bool count, unsigned long long* d_counts,
                          hipStream_t stream, KernelTimer* timer);
// The batch's bounce loop (capturable) and its resolve / ray-count epilogue.
hipError_t launchMptBatchBounces(const MptArgs& a, const TraceScratch& ts, bool count, unsigned long long* d_counts,
                                 hipStream_t stream, KernelTimer* timer);
hipError_t launchMptBatchEpilogue(const MptArgs& a, unsigned long long* d_counts, hipStream_t stream);
hipError_t launchMptFrame(const SceneDev& s, const MptArgs& a, const TraceScratch& ts, float4* out, bool count,
                          unsigned long long* d_counts, hipStream_t stream, KernelTimer* timer);
// Pieces used with graph capture: the bounce loop (capturable, no host-varying arguments) and
// the per-call prologue/epilogue launched outside the graph.
hipError_t launchMptBounces(const SceneDev& s, const MptArgs& a, const TraceScratch& ts, bool count,
                            unsigned long long* d_counts, hipStream_t stream, KernelTimer* timer);
hipError_t launchMptPrologue(const MptArgs& a, hipStream_t stream);
hipError_t launchMptEpilogue(const SceneDev& s, const MptArgs& a, float4* out, unsigned long long* d_counts,
                             hipStream_t stream);

// WARDiffPathTracer primal and adjoint (megakernel, one path per thread).
// One view of a batched WARDiffPathTracer launch: its camera, frame seed base, dL/dI image
// (3 floats per pixel) and output image (RGBA). Every view of a batch has the launch's W, H, spp.
struct WarView {
    CameraDev cam;
    uint32_t frame;
    const float* dLdI;
    float4* out;
};
struct WarArgs {
    uint32_t W, H, spp, frame, maxBounces;
    uint32_t useBSDFSampling, useNEE, useMIS, useAntithetic;
    float4* out;            // primal image (RGBA)
    float4* tmp;            // per-sample radiance (W * H * spp)
    const float* dLdI;      // adjoint input (3 floats per pixel)
    float* gradMat;         // material gradients [hashSize][nMat * 20]
    uint32_t hashSize, matGradDim;
    float* gradTex;         // per-texel baseColor gradients (texel layout), or null
    unsigned long long* rays;  // [closest, visibility] counters, or null
    // Gradient-image debug modes (DiffDebugParams, SharedTypes.slang:98-104): material dbgMat's
    // parameter entry at dbgOffset of the 20-float layout, direction dbgGrad; per-sample
    // derivatives beyond dbgThreshold in magnitude are dropped (assertThreshold, Params.slang:61).
    uint32_t dbgMat, dbgOffset;
    float dbgGrad[4];
    float dbgThreshold;
    float4* dColor;         // debug image (RGBA), or null
    // Batched views (null: one view with the scene's camera, frame, dLdI and out above). Path id
    // v * W * H * spp + p is path p of view v; tmp holds every view's samples in that order.
    const WarView* views;
    uint32_t nViews;
    // Adjoint and debug-image kernels: per-thread hit log of the first pass, read by the replay
    // (logDepth words per thread of the persistent grid, [entry][thread]); null: the replay traces.
    uint32_t* hitLog;
    uint32_t logDepth;
};
hipError_t launchWarPrimal(const SceneDev& s, const WarArgs& a, const TraceScratch& ts, hipStream_t stream,
                           KernelTimer* timer);
hipError_t launchWarAdjoint(const SceneDev& s, const WarArgs& a, const TraceScratch& ts, hipStream_t stream,
                            KernelTimer* timer);
// ForwardDiffDebug for a material variable: the primal image into out and, into dColor, the red
// channel's derivative along the debug direction (red, 0, 0, 1), both averaged over spp; tmp holds
// W * H * spp per-sample (radiance, derivative).
hipError_t launchWarDebugImage(const SceneDev& s, const WarArgs& a, const TraceScratch& ts, hipStream_t stream,
                               KernelTimer* timer);
// The adjoint (debug = false) or ForwardDiffDebug sample kernel (debug = true) instances with
// the transmission-lobe adjoint, on `blocks` blocks of ts's configuration.
hipError_t launchWarTransKernel(const SceneDev& s, const WarArgs& a, const TraceScratch& ts, hipStream_t stream,
                                uint32_t blocks, bool debug);
// Geometry gradients of WARDiffPathTracer (warped-area reparameterization, maxBounces = 0).
struct GeomArgs {
    uint32_t W, H, spp, frame;
    uint32_t useBSDFSampling, useNEE, useMIS, useAntithetic;
    uint32_t aux;              // auxiliary rays per reparameterization
    float kappa, kappaScreen;  // vMF concentrations (10^Log10vMFConcentration[Screen])
    uint32_t neeHeld;          // 1: NEE direction held, no area factor (useSolidAngleNEEWarp); 0: reference
    const WarView* views;      // batched views (camera, frame, dLdI per view), or null
    uint32_t nViews;
    const float* dLdI;
    float* gradPos;            // MeshPosition gradients [hashSize][3 * meshVerts]
    float* gradNrm;            // MeshNormal gradients, same layout
    uint32_t hashSize, meshV0, meshVerts;
    unsigned long long* rays;  // closest-hit ray counter (primary, scatter and auxiliary rays), or null
    // Aux-hit log: 3 * aux words per thread of the persistent grid ([word][thread]), so each
    // warp's second loop over its aux rays reads the first loop's hits; null: both loops trace.
    uint32_t* auxLog;
    uint32_t maxBounces;
    // Vertex records of the forward pass: kGeomVertexWords * (maxBounces + 2) words per thread of
    // the persistent grid ([word][thread]), read back by the reverse sweep.
    uint32_t* vtxLog;
};
constexpr uint32_t kGeomVertexWords = 17;
hipError_t launchWarGeom(const SceneDev& s, const GeomArgs& a, const TraceScratch& ts, hipStream_t stream, KernelTimer* timer);
hipError_t launchAggregate(const float* tmp, uint32_t hashSize, uint32_t dim, float* out, hipStream_t stream);
hipError_t launchFill(float* p, uint64_t n, float v, hipStream_t stream);
hipError_t launchFillRGBA(float4* p, uint64_t n, float r, float g, float b, float a, hipStream_t stream);

// BVH refit after vertex updates of one mesh (fhp_refit.hip).
struct RefitArgs {
    float4* vtx;             // vertex records of the scene
    float4* triPos;          // leaf-order triangle records
    float* bvh;              // BVH4 nodes (16 words each, packBVH4Node)
    const uint32_t* idx;     // triangle vertex indices
    const float* pos;        // new positions of vertices [v0, v0 + nVerts) (3 floats each, device);
                             // null when the vertex records already hold them
    const float* nrm;        // new normals (or null)
    uint32_t v0, nVerts, nTris;
    const uint32_t* levelNodes;  // device: node ids by depth
    const uint32_t* levelOff;    // host: nLevels + 1 offsets into levelNodes
    uint32_t nLevels;
    float absPad;
};
hipError_t launchRefit(const RefitArgs& r, hipStream_t stream);

// Node transform of one mesh (Scene::updateNodeTransform + the animation update): vertex records
// [v0, v0 + nVerts) become M * rest position and N * rest normal (M: 3x4 row-major affine, N: the
// inverse transpose of its 3x3 part).
struct NodeXformArgs {
    float4* vtx;
    const float4* rest;  // rest-pose vertex records, same layout as vtx
    uint32_t v0, nVerts;
    float M[12], N[9];
};
hipError_t launchNodeTransform(const NodeXformArgs& a, hipStream_t stream);
// LightCollection::updateTrianglePositions (LightCollection.cpp:550-574, UpdateTriangleVertices.cs.slang)
// then LightBVH::refit (LightBVH.cpp:49-89, LightBVHRefit.cs.slang): every emissive triangle's
// vertices, face normal and area from the vertex records (flux kept), then the light BVH's boxes
// and normal cones, leaves first, then internal nodes from the deepest level up.
struct LightRefitArgs {
    float* lights;              // 20 floats per light
    uint32_t nLights;
    const float4* vtx;
    const uint32_t* idx;
    uint32_t* lbvh;             // packed nodes, 8 words each, then their sampling records
    uint32_t nNodes;
    const uint32_t* lbvhTri;
    const uint32_t* nodes;      // device: refit schedule
    const uint32_t* off;        // host: depth + 2 offsets (internal nodes by depth, then leaves)
    uint32_t depth;
};
hipError_t launchLightRefit(const LightRefitArgs& a, hipStream_t stream);
// Env map importance map (fhp_envmap.hip): dim x dim finest level (power of two) with
// samplesPerAxis^2 luminance samples per texel, all levels into out ((4 dim^2 - 1) / 3 floats).
hipError_t launchEnvImportance(const SceneDev& s, float* out, uint32_t dim, uint32_t samplesPerAxis, hipStream_t st);
hipError_t launchEnvSampleBatch(const SceneDev& s, const float* u, uint32_t n, float* out, hipStream_t st);
hipError_t launchEnvPdfBatch(const SceneDev& s, const float* dirs, uint32_t n, float* pdf, hipStream_t st);

// AccumulatePass (Accumulate.cs.slang, AccumulatePass.cpp). Precision modes use the values of
// AccumulatePass::Precision (AccumulatePass.h:70-75); input formats those of _INPUT_FORMAT
// (Accumulate.cs.slang:40-43).
constexpr int kAccumDouble = 0, kAccumSingle = 1, kAccumSingleCompensated = 2;
constexpr int kAccumInputFloat = 0, kAccumInputUint = 1, kAccumInputSint = 2;
constexpr uint32_t kAccumEnabled = 1, kAccumMovingAverage = 2, kAccumFirst = 4;
struct AccumArgs {
    const void* in;   // n RGBA texels (float, uint32 or int32)
    float4* out;      // n RGBA float
    float4* sum;      // Single / SingleCompensated running sum (or moving average)
    float4* corr;     // SingleCompensated correction term
    double2* sumD;    // Double: 2 double2 per pixel
    uint64_t n;
    uint32_t count;   // frames accumulated before this one
    uint32_t flags;   // kAccum* flags
    int mode, inputFormat;
};
hipError_t launchAccumulate(const AccumArgs& a, hipStream_t stream);

// ImageCompare metrics in the order of its errorMetrics table (ImageCompare.cpp:256-261).
constexpr int kMetricMSE = 0, kMetricRMSE = 1, kMetricMAE = 2, kMetricMAPE = 3;
constexpr uint32_t kComparePartialsMax = 2048;
struct CompareResult {
    double sum;       // sum of the per-pixel errors
    float minErr, maxErr;  // min / max of the float per-pixel errors (NaN left out)
};
uint32_t comparePartials(uint64_t n);
// part: comparePartials(n) entries; res: one entry. errMap may be null.
hipError_t launchCompare(int metric, const float4* a, const float4* b, uint64_t n, int nch, float* errMap,
                         CompareResult* part, CompareResult* res, hipStream_t stream);
hipError_t launchHeatMap(const float* err, uint64_t n, const CompareResult* res, float4* out, hipStream_t stream);

int numCUs();

}  // namespace fhp
