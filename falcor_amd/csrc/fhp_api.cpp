// C ABI of libfhp (declared in include/fhp.h): scenes, ray queries, render passes, gradients.
#include "../../include/fhp.h"
#include "fhp_internal.h"
#include "scene_build.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <vector>

namespace {
using namespace fhp;

thread_local std::string g_error;
thread_local std::string g_warning;

int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}
#define HIP_OR_FAIL(expr)                                                                         \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) return fail(FHP_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Device buffer owned by the library.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t alloc(size_t n) {
        if (n <= bytes && p) return hipSuccess;
        release();
        if (n == 0) return hipSuccess;
        hipError_t e = hipMalloc(&p, n);
        if (e == hipSuccess) bytes = n;
        return e;
    }
    template <typename T> hipError_t upload(const std::vector<T>& v) {
        hipError_t e = alloc(std::max<size_t>(v.size() * sizeof(T), 16));
        if (e != hipSuccess || v.empty()) return e;
        return hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
    }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};
}  // namespace

struct FhpSession_ {
    int device;
    hipStream_t stream;
};

struct FhpScene_ {
    HostScene h;
    DevBuf bvh, triPos, idx, vtx, triMat, mats, texels, texDesc, lights, active, triToLight, lbvh, lbvhTri, lbvhMask;
    DevBuf meshTri;  // first triangle of each mesh (PackedHitInfo instance ids)
    DevBuf levels;   // BVH refit schedule (node ids by depth)
    DevBuf envTex;   // environment map texels (RGB fp32)
    DevBuf envImp;   // env map importance levels (EnvMapSampler)
    DevBuf lrefit;   // light BVH refit schedule
    DevBuf alights;  // analytic lights
    DevBuf restVtx;  // rest-pose vertex records for node transforms (uploaded on first use)
    bool restValid = false;
    uint32_t leafSize = 4;  // BVH4 leaf size (triangles)
    DevBuf ovf;  // traversal overflow stack
    DevBuf err;  // trace error word (kTraceErr* bits set by the traversal kernels)
    DevBuf matsS, texelsS;  // frozen sampling parameters (finite-difference mode)
    TraceScratch ts{};      // default traversal configuration (ray queries)
    uint32_t stackNeed = 0; // traversal stack entries a ray can need (from the BVH build)
    SceneDev dev{};
    uint64_t deviceBytes = 0;
};

namespace fhp {
// Camera jitter patterns of VBufferRT (GBufferBase.cpp:202-218): none (Center), the 8 standard
// Direct3D sample positions (DxSamplePattern.cpp), Halton bases 2 and 3
// (HaltonSamplePattern.cpp), or stratified bins visited in a shuffled order per round
// (StratifiedSamplePattern.cpp: std::mt19937 with its default seed, std::shuffle and
// std::uniform_real_distribution<float>, so a libstdc++ build reproduces the reference's sequence).
struct JitterPattern {
    int kind = 0;  // 0 Center, 1 DirectX, 2 Halton, 3 Stratified
    uint32_t count = 0, cur = 0, binsX = 1, binsY = 1;
    std::vector<uint32_t> perm;
    std::mt19937 rng;

    // Returns the sample count the pattern uses (the reference stores it back into sampleCount).
    uint32_t create(int k, uint32_t sampleCount) {
        kind = k;
        cur = 0;
        rng = std::mt19937();
        switch (k) {
        case 1: count = 8; break;
        case 2: count = sampleCount; break;
        case 3: {
            count = std::min(std::max(sampleCount, 1u), 1024u);
            binsX = uint32_t(std::sqrt(double(count)));
            binsY = count / binsX;
            while (binsX * binsY != count) {
                binsX++;
                binsY = count / binsX;
            }
            perm.resize(count);
            for (uint32_t i = 0; i < count; i++) perm[i] = i;
            break;
        }
        default: count = sampleCount; break;
        }
        return count;
    }
    static float halton(uint32_t index, uint32_t base) {
        float result = 0.f, factor = 1.f;
        for (; index > 0; index /= base) {
            factor /= float(base);
            result += factor * float(index % base);
        }
        return result;
    }
    // Next jitter sample in [-0.5, 0.5)^2 (pixel units), or (0, 0) for Center.
    void next(float& x, float& y) {
        x = y = 0.f;
        if (kind == 1) {
            static const float kDx[8][2] = {{1.f / 16.f, -3.f / 16.f}, {-1.f / 16.f, 3.f / 16.f}, {5.f / 16.f, 1.f / 16.f},
                                            {-3.f / 16.f, -5.f / 16.f}, {-5.f / 16.f, 5.f / 16.f}, {-7.f / 16.f, -1.f / 16.f},
                                            {3.f / 16.f, 7.f / 16.f}, {7.f / 16.f, -7.f / 16.f}};
            x = kDx[cur % 8][0];
            y = kDx[cur % 8][1];
            cur++;
        } else if (kind == 2) {
            const float hx = halton(cur, 2) + 0.5f, hy = halton(cur, 3) + 0.5f;
            cur++;
            if (count != 0) cur %= count;
            x = (hx - std::floor(hx)) - 0.5f;
            y = (hy - std::floor(hy)) - 0.5f;
        } else if (kind == 3) {
            std::uniform_real_distribution<float> dist;
            if (cur == 0) std::shuffle(perm.begin(), perm.end(), rng);
            const uint32_t bin = perm[cur];
            const uint32_t i = bin % binsX, j = bin / binsX;
            cur = (cur + 1) % count;
            const float ux = dist(rng);
            const float px = (float(i) + ux) / float(binsX);
            const float uy = dist(rng);
            const float py = (float(j) + uy) / float(binsY);
            x = px - 0.5f;
            y = py - 0.5f;
        }
    }
};
}  // namespace fhp

struct FhpPass_ {
    std::string kind;
    std::map<std::string, double> props;
    // Wavefront buffers (forward pass).
    DevBuf thp, prevN, rng, Lp, Ls, rayQ0, rayQ1, pathQ0, pathQ1, hitQ, hitIdx, shQ, pixHit, counters, stats;
    DevBuf viewTab, sceneTab;  // batch mode tables
    DevBuf pathView;           // batch mode: view index per path
    DevBuf warViewTab;         // batched WARDiffPathTracer views
    DevBuf warLog;             // WARDiffPathTracer adjoint hit log
    DevBuf auxLog;             // WARDiffPathTracer vertex-gradient aux-hit log
    DevBuf vtxLog;             // WARDiffPathTracer vertex-gradient path vertex records
    uint32_t cap = 0, capPix = 0;
    KernelTimer timer;
    // AccumulatePass state: running sums of the current precision mode, frames accumulated,
    // frame shape (AccumulatePass.h:130-150).
    DevBuf accSum, accCorr;
    uint32_t accCount = 0, accW = 0, accH = 0;
    int accMode = -1;  // precision mode the sums are laid out for (-1: none yet)
    // VBufferRT camera jitter pattern (created from samplePattern / sampleCount).
    JitterPattern jitter;
    // Bound input channels of MinimalPathTracer (MinimalPathTracer.cpp:48-53): "vbuffer", "viewW".
    const uint32_t* inVbuf = nullptr;
    const float4* inViewW = nullptr;
    // WARDiffPathTracer: bound dColor output (WARDiffPathTracer.cpp:50) and the gradient-image
    // debug parameters (setDiffDebugParams, WARDiffPathTracer.cpp:689-695).
    float4* outDColor = nullptr;
    int dbgVar = 0;
    uint32_t dbgId = 0, dbgOffset = 0;
    float dbgGrad[4] = {0.f, 0.f, 0.f, 0.f};
    // Captured bounce loop of the forward pass (property useGraph), keyed by everything the
    // captured kernels read by value.
    hipGraphExec_t graph = nullptr;
    hipStream_t capStream = nullptr;
    std::vector<uint8_t> graphKey;
    // Captured bounce loop of the batch mode, keyed by its arguments (the view and scene tables
    // are read from device memory, so new cameras, frames and outputs replay the same graph).
    hipGraphExec_t batchGraph = nullptr;
    std::vector<uint8_t> batchKey;
    // Pinned staging of the batch tables; stageDone marks the end of the last copy out of it.
    void* stage = nullptr;
    size_t stageBytes = 0;
    hipEvent_t stageDone = nullptr;
    void dropGraph() {
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
        graphKey.clear();
        if (batchGraph) (void)hipGraphExecDestroy(batchGraph);
        batchGraph = nullptr;
        batchKey.clear();
    }
    ~FhpPass_() {
        if (timer.created)
            for (auto& e : timer.ev) (void)hipEventDestroy(e);
        dropGraph();
        if (capStream) (void)hipStreamDestroy(capStream);
        if (stageDone) {
            (void)hipEventSynchronize(stageDone);
            (void)hipEventDestroy(stageDone);
        }
        if (stage) (void)hipHostFree(stage);
    }
    double get(const char* k) const {
        auto it = props.find(k);
        return it == props.end() ? 0.0 : it->second;
    }
};

struct FhpGrads_ {
    FhpScene scene = nullptr;
    uint32_t hashSize = 1, dim = 0;
    DevBuf tmp, agg, tex;
    bool withTex = false;
    // MeshPosition / MeshNormal gradients of one mesh: hashSize rows of 3 * meshVerts floats each.
    bool withMesh = false;
    uint32_t mesh = 0, meshV0 = 0, meshVerts = 0;
    DevBuf posTmp, nrmTmp, posAgg, nrmAgg;
};

namespace {
// Property tables: names and defaults of the reference passes.
const std::map<std::string, double>& mptDefaults() {
    // MinimalPathTracer.h:79-83; useEmissiveNEE/useEnvNEE/emissiveSampler/countTraversal are extensions.
    static const std::map<std::string, double> d = {{"maxBounces", 3}, {"computeDirect", 1}, {"useImportanceSampling", 1},
                                                   {"useEmissiveNEE", 0}, {"useEnvNEE", 0}, {"emissiveSampler", 1},
                                                   {"countTraversal", 0},
                                                   {"timeKernels", 0}, {"useGraph", 1}, {"traceBlockSize", kDefaultBlock},
                                                   {"traceStackDepth", kDefaultLds}, {"traceBlocksPerCU", kDefaultBlocksPerCU},
                                                   {"splitVisibilityLaunch", 0}, {"tailPaths", 65536},
                                                   {"tailLanes", 8}};
    return d;
}
// WARDiffPathTracer's DiffMode (WARDiffPathTracer.h:56-62) and DiffVariableType (SharedTypes.slang:78-84).
constexpr int kDiffPrimal = 0, kDiffBackward = 1, kDiffForwardDebug = 2, kDiffBackwardDebug = 3;
// Deepest path of the vertex-gradient kernel (its per-thread vertex records grow with it).
constexpr int kMaxGeomBounces = 8;
constexpr int kDiffVarNone = 0, kDiffVarMaterial = 1, kDiffVarGeometryTranslation = 2;
const std::map<std::string, double>& warDefaults() {
    // WARDiffPathTracer.h:120-170 (StaticParams) and Params.slang.
    static const std::map<std::string, double> d = {
        {"samplesPerPixel", 1}, {"maxBounces", 0},        {"diffMode", kDiffForwardDebug}, {"sampleGenerator", 0},
        {"diffVarName", 0},     {"assertThreshold", 1e9},
        {"fixedSeed", 0},       {"useFixedSeed", 0},      {"useBSDFSampling", 1},   {"useNEE", 1},
        {"useMIS", 1},          {"useWAR", 0},            {"auxSampleCount", 16},   {"Log10vMFConcentration", 5},
        {"Log10vMFConcentrationScreen", 5},               {"boundaryTermBeta", 0.01}, {"useAntitheticSampling", 1},
        {"useSolidAngleNEEWarp", 0},  // extension: NEE vertex gradients with the sampled direction held
        {"useHitLog", 1},  // extension: adjoint replays and second aux loops read the first pass's hits
        {"countTraversal", 0},  {"timeKernels", 0},       {"traceBlockSize", 128},
        {"traceStackDepth", 24},                          {"traceBlocksPerCU", 10}};  // 1280 threads per CU: 5 waves per SIMD
    return d;
}

const std::map<std::string, double>& accDefaults() {
    // AccumulatePass.h:135-145: enabled, autoReset, precisionMode (Single), maxFrameCount (0: no
    // limit), overflowMode (Stop).
    static const std::map<std::string, double> d = {
        {"enabled", 1}, {"autoReset", 1}, {"precisionMode", kAccumSingle}, {"maxFrameCount", 0}, {"overflowMode", 0}};
    return d;
}
const std::map<std::string, double>& vbufDefaults() {
    // GBufferBase.h:96-106 and VBufferRT.h:74-76: samplePattern (Center), sampleCount,
    // useAlphaTest, adjustShadingNormals, forceCullMode, cull (Back), useTraceRayInline, useDOF,
    // plus the traversal launch properties.
    static const std::map<std::string, double> d = {
        {"samplePattern", 0}, {"sampleCount", 16}, {"useAlphaTest", 1}, {"adjustShadingNormals", 1},
        {"forceCullMode", 0}, {"cull", 2}, {"useTraceRayInline", 0}, {"useDOF", 1},
        {"traceBlockSize", kDefaultBlock}, {"traceStackDepth", kDefaultLds}, {"traceBlocksPerCU", kDefaultBlocksPerCU}};
    return d;
}
// VBufferRT property values this build renders: the four sample patterns and no forced face
// culling.
std::string vbufCheck(const std::string& k, double v, const std::map<std::string, double>& props) {
    if (k == "samplePattern" && !(v == 0 || v == 1 || v == 2 || v == 3))
        return "samplePattern: 0 Center, 1 DirectX, 2 Halton, 3 Stratified";
    if (k == "sampleCount" && !(v >= 0 && v <= 4294967295.0 && double(uint64_t(v)) == v))
        return "sampleCount: a non-negative integer";
    const double force = k == "forceCullMode" ? v : props.at("forceCullMode");
    const double cull = k == "cull" ? v : props.at("cull");
    if ((k == "forceCullMode" || k == "cull") && force != 0 && cull != 0)
        return "forceCullMode with cull Front/Back is not supported (rays hit both faces)";
    return "";
}
// Validates one AccumulatePass property value; returns an error message or "".
std::string accCheck(const std::string& k, double v) {
    const bool integral = v >= 0 && v <= 4294967295.0 && double(uint64_t(v)) == v;
    if (k == "precisionMode" && !(integral && v <= 2)) return "precisionMode: 0 Double, 1 Single, 2 SingleCompensated";
    if (k == "overflowMode" && !(integral && v <= 2)) return "overflowMode: 0 Stop, 1 Reset, 2 EMA";
    if (k == "maxFrameCount" && !integral) return "maxFrameCount: a non-negative integer";
    return "";
}

constexpr uint32_t kMinLds = 16;
constexpr uint32_t kMaxThreadsPerCU = 2048;  // 32 resident waves per CU

TraceScratch traceConfig(const FhpScene_* s, uint32_t block, uint32_t lds, uint32_t blocksPerCU) {
    TraceScratch t{};
    t.block = block;
    t.lds = lds;
    t.gridBlocks = uint32_t(numCUs()) * blocksPerCU;
    t.ovfDepth = s->stackNeed > lds ? s->stackNeed - lds : 0u;
    t.ovf = s->ovf.as<uint32_t>();
    t.err = s->err.as<uint32_t>();
    return t;
}
// Validates the traversal launch properties (traceBlockSize, traceStackDepth, traceBlocksPerCU).
bool validLaunch(double block, double lds, double bpc) {
    if (!validTraceConfig(uint32_t(block), uint32_t(lds)) || double(uint32_t(block)) != block || double(uint32_t(lds)) != lds)
        return false;
    return bpc >= 1 && double(uint32_t(bpc)) == bpc && uint32_t(bpc) * uint32_t(block) <= kMaxThreadsPerCU;
}
// Launch configuration of a pass on scene s. The persistent grid indexes the overflow stack as
// [entry][global thread] over gridBlocks * block threads and ovfDepth entries: refuse a
// configuration whose indices would not fit the scene's overflow buffer.
int passTrace(const FhpScene_* s, const FhpPass_* p, TraceScratch& t) {
    t = traceConfig(s, uint32_t(p->get("traceBlockSize")), uint32_t(p->get("traceStackDepth")),
                    uint32_t(p->get("traceBlocksPerCU")));
    const uint64_t words = uint64_t(t.ovfDepth) * t.gridBlocks * t.block;
    if (words && (!s->ovf.p || words * 4 > s->ovf.bytes))
        return fail(FHP_ESTATE, "traversal overflow stack smaller than the launch needs");
    return FHP_OK;
}

// Hit log of the WARDiffPathTracer adjoint and debug-image kernels (useHitLog): one word per
// closest-hit trace of a path (at most maxBounces + 2) for every thread of the persistent grid,
// so the replay pass traces nothing. Triangle ids must leave bit 31 free for the visibility bit.
int prepareHitLog(FhpPass_* p, const FhpScene_* s, const TraceScratch& ts, WarArgs& a) {
    a.hitLog = nullptr;
    a.logDepth = 0;
    const uint64_t depth = uint64_t(a.maxBounces) + 2;
    if (p->get("useHitLog") == 0 || depth > 64 || s->h.nT >= 0x7fffffffu) return FHP_OK;
    HIP_OR_FAIL(p->warLog.alloc(depth * ts.gridBlocks * ts.block * 4));
    a.hitLog = p->warLog.as<uint32_t>();
    a.logDepth = uint32_t(depth);
    return FHP_OK;
}

// The debug parameters the reference's constructor sets for its two named debug variables
// (WARDiffPathTracer.cpp:84-93): CBOX_BUNNY_MATERIAL = baseColor of material 0 along (1, 1, 1),
// CBOX_BUNNY_TRANSLATION = translation of mesh 0 along +y.
void applyDiffVarName(FhpPass_* p) {
    const int name = int(p->get("diffVarName"));
    if (name == 1) {
        p->dbgVar = kDiffVarMaterial;
        p->dbgId = 0;
        p->dbgOffset = 0;
        const float g[4] = {1.f, 1.f, 1.f, 0.f};
        std::memcpy(p->dbgGrad, g, sizeof(g));
    } else if (name == 2) {
        p->dbgVar = kDiffVarGeometryTranslation;
        p->dbgId = 0;
        p->dbgOffset = 0;
        const float g[4] = {0.f, 1.f, 0.f, 0.f};
        std::memcpy(p->dbgGrad, g, sizeof(g));
    }
}

int ensurePassBuffersN(FhpPass p, uint64_t n, uint64_t npix);
int ensurePassBuffers(FhpPass p, uint32_t W, uint32_t H, uint32_t spp) {
    return ensurePassBuffersN(p, uint64_t(W) * H * spp, uint64_t(W) * H);
}
// Wavefront buffers for n paths over npix pixels.
int ensurePassBuffersN(FhpPass p, uint64_t n, uint64_t npix) {
    if (n > 0xffffffffull / 3) return fail(FHP_EINVAL, "frame too large");
    if (n <= p->cap && npix <= p->capPix) return FHP_OK;
    p->dropGraph();
    HIP_OR_FAIL(p->thp.alloc(n * 16));
    HIP_OR_FAIL(p->prevN.alloc(n * 16));
    HIP_OR_FAIL(p->rng.alloc(n * 16));
    HIP_OR_FAIL(p->Lp.alloc(n * 16));
    HIP_OR_FAIL(p->Ls.alloc(n * 16));
    HIP_OR_FAIL(p->rayQ0.alloc(n * 32));
    HIP_OR_FAIL(p->rayQ1.alloc(n * 32));
    HIP_OR_FAIL(p->pathQ0.alloc(n * 4));
    HIP_OR_FAIL(p->pathQ1.alloc(n * 4));
    HIP_OR_FAIL(p->hitQ.alloc(n * 16));
    HIP_OR_FAIL(p->shQ.alloc(n * 48));
    HIP_OR_FAIL(p->pixHit.alloc(npix * 16));
    HIP_OR_FAIL(p->counters.alloc(kCounterWords * 4));
    HIP_OR_FAIL(p->hitIdx.alloc(n * 4));
    HIP_OR_FAIL(p->stats.alloc(64));
    p->cap = uint32_t(n);
    p->capPix = uint32_t(npix);
    return FHP_OK;
}
}  // namespace

// MinimalPathTracer arguments of a frame from the pass properties and the scene (everything but
// the wavefront buffers). Returns FHP_OK or an argument error.
int mptFrameArgs(const FhpPass_* p, const FhpScene_* s, uint32_t W, uint32_t H, uint32_t spp, uint32_t frame, MptArgs& a) {
    std::memset(&a, 0, sizeof(a));  // padding too: the bytes key the captured graph
    a.W = W; a.H = H; a.spp = spp; a.frame = frame;
    a.maxBounces = uint32_t(p->get("maxBounces"));
    a.computeDirect = p->get("computeDirect") != 0;
    a.importance = p->get("useImportanceSampling") != 0;
    a.nee = p->get("useEmissiveNEE") != 0;
    a.lightKind = uint32_t(p->get("emissiveSampler"));
    a.separateShadow = p->get("splitVisibilityLaunch") != 0;
    a.tailPaths = uint32_t(p->get("tailPaths"));
    a.tailLanes = uint32_t(p->get("tailLanes"));
    a.envNee = p->get("useEnvNEE") != 0;
    if (a.nee && !a.computeDirect) return fail(FHP_EINVAL, "useEmissiveNEE requires computeDirect");
    if (a.envNee && !a.computeDirect) return fail(FHP_EINVAL, "useEnvNEE requires computeDirect");
    const bool envLight = a.envNee && s->dev.envImp && (s->dev.envFlags & kEnvLight);
    a.shadowRays = (a.nee || s->dev.nALights || envLight) ? 1u : 0u;
    a.lightMode = (s->dev.nALights || envLight) ? 2u : (a.nee ? 1u : 0u);
    a.nPaths = W * H * spp;
    a.nPix = W * H;
    if (p->inVbuf && !s->h.meshTriContiguous)
        return fail(FHP_ESTATE, "a V-buffer input needs each mesh's triangles to be contiguous");
    a.vbuf = p->inVbuf;
    a.viewW = p->inViewW;
    a.keepMisses = s->dev.envTex && (s->dev.envFlags & kEnvLight);
    return FHP_OK;
}

int prepareTimer(FhpPass p) {
    p->timer.on = p->get("timeKernels") != 0;
    if (p->timer.on && !p->timer.created) {
        for (auto& e : p->timer.ev) HIP_OR_FAIL(hipEventCreate(&e));
        p->timer.created = true;
    }
    p->timer.n = 0;
    return FHP_OK;
}
int collectTimer(FhpPass p, hipStream_t st) {
    if (!p->timer.on || p->timer.n == 0) return FHP_OK;
    HIP_OR_FAIL(hipStreamSynchronize(st));
    for (int i = 0; i < p->timer.n; i++) {
        float ms = 0.f;
        HIP_OR_FAIL(hipEventElapsedTime(&ms, p->timer.ev[2 * i], p->timer.ev[2 * i + 1]));
        p->timer.ms[p->timer.kind[i]] += ms;
        p->timer.launches[p->timer.kind[i]]++;
    }
    p->timer.n = 0;
    return FHP_OK;
}

// Host-array scene mutators are synchronous: before they overwrite or reallocate buffers that
// kernels read, every queued launch on every stream of the device has finished (a caller may
// render on any stream, the session's non-blocking one included).
int waitForDevice() {
    HIP_OR_FAIL(hipDeviceSynchronize());
    return FHP_OK;
}

// Builds the BVH, light collection and light BVH of s->h and uploads everything the kernels
// read. Called at creation and after vertex updates (buffers are reused when large enough; a
// frozen sampling copy of the materials stays in place).
int buildAndUpload(FhpScene_* s) {
    HostScene& h = s->h;
    if (int rc = waitForDevice()) return rc;
    try {
        buildBVH4(h, s->leafSize);
        buildLights(h);
        buildLightBVH(h);
    } catch (const std::exception& e) {
        return fail(FHP_EBUILD, e.what());
    }
    const std::vector<float> vtx = packVertexRecords(h);
    HIP_OR_FAIL(s->bvh.upload(h.bvh));
    HIP_OR_FAIL(s->triPos.upload(h.triPos));
    HIP_OR_FAIL(s->idx.upload(h.idx));
    HIP_OR_FAIL(s->vtx.upload(vtx));
    HIP_OR_FAIL(s->triMat.upload(h.triMat));
    HIP_OR_FAIL(s->mats.upload(h.mats));
    HIP_OR_FAIL(s->texels.upload(h.texels));
    HIP_OR_FAIL(s->texDesc.upload(h.texDesc));
    HIP_OR_FAIL(s->lights.upload(h.lights));
    HIP_OR_FAIL(s->active.upload(h.activeLights));
    HIP_OR_FAIL(s->triToLight.upload(h.triToLight));
    HIP_OR_FAIL(s->lbvh.upload(h.lbvhPacked));
    HIP_OR_FAIL(s->lbvhTri.upload(h.lbvhTri));
    HIP_OR_FAIL(s->lbvhMask.upload(h.lbvhMask));
    HIP_OR_FAIL(s->meshTri.upload(h.meshTriOff));
    HIP_OR_FAIL(s->levels.upload(h.levelNodes));
    HIP_OR_FAIL(s->lrefit.upload(h.lrefitNodes));
    // Overflow stack for the persistent traversal grid, sized for the smallest LDS stack and the
    // largest grid any configuration launches.
    s->stackNeed = h.stackNeed + 1;
    const uint32_t worst = s->stackNeed > kMinLds ? s->stackNeed - kMinLds : 0u;
    if (worst) HIP_OR_FAIL(s->ovf.alloc(size_t(worst) * uint32_t(numCUs()) * kMaxThreadsPerCU * 4));
    if (!s->err.p) {
        HIP_OR_FAIL(s->err.alloc(16));
        HIP_OR_FAIL(hipMemset(s->err.p, 0, 16));
    }
    s->ts = traceConfig(s, kDefaultBlock, kDefaultLds, kDefaultBlocksPerCU);

    SceneDev& dv = s->dev;
    const bool frozen = s->matsS.p != nullptr && dv.matsS == s->matsS.as<float>();
    dv.bvh = s->bvh.as<float4>();
    dv.triPos = s->triPos.as<float4>();
    dv.idx = s->idx.as<uint32_t>();
    dv.vtx = s->vtx.as<float4>();
    dv.triMat = s->triMat.as<uint32_t>();
    dv.mats = s->mats.as<float>();
    dv.texels = s->texels.as<float>();
    if (!frozen) {
        dv.matsS = dv.mats;
        dv.texelsS = dv.texels;
    }
    dv.texDesc = s->texDesc.as<int4>();
    dv.lights = s->lights.as<float4>();
    dv.activeLights = s->active.as<uint32_t>();
    dv.triToLight = s->triToLight.as<uint32_t>();
    dv.lbvh = s->lbvh.as<uint4>();
    dv.lbvhS = reinterpret_cast<const float4*>(s->lbvh.as<uint32_t>() + 8 * size_t(h.nLbvh));
    dv.lbvhTri = s->lbvhTri.as<uint32_t>();
    dv.lbvhMask = s->lbvhMask.as<uint2>();
    dv.cam.pos = mk3(h.cam[0], h.cam[1], h.cam[2]);
    dv.cam.U = mk3(h.camU[0], h.camU[1], h.camU[2]);
    dv.cam.V = mk3(h.camV[0], h.camV[1], h.camV[2]);
    dv.cam.W = mk3(h.camW[0], h.camW[1], h.camW[2]);
    dv.cam.nearZ = h.cam[12];
    dv.cam.farZ = h.cam[13];
    dv.meshTriOff = h.meshTriContiguous ? s->meshTri.as<uint32_t>() : nullptr;
    dv.nMeshes = uint32_t(h.meshOff.size() - 1);
    dv.hit = HitInfoLayout{h.hitCompressed, h.hitTypeBits, h.hitInstBits, h.hitPrimBits};
    dv.viewProjZ = make_float4(h.viewProj[8], h.viewProj[9], h.viewProj[10], h.viewProj[11]);
    dv.viewProjW = make_float4(h.viewProj[12], h.viewProj[13], h.viewProj[14], h.viewProj[15]);
    dv.nTris = h.nT;
    dv.hasTransmission = hasTransmission(h) ? 1u : 0u;
    dv.nLights = h.nLights;
    dv.nActive = uint32_t(h.activeLights.size());
    dv.nLbvh = h.nLbvh;
    s->deviceBytes = 0;
    for (DevBuf* b : {&s->bvh, &s->triPos, &s->idx, &s->vtx, &s->triMat, &s->mats, &s->texels, &s->texDesc, &s->lights,
                      &s->active, &s->triToLight, &s->lbvh, &s->lbvhTri, &s->lbvhMask, &s->ovf, &s->matsS, &s->texelsS,
                      &s->lrefit, &s->restVtx})
        s->deviceBytes += b->bytes;
    return FHP_OK;
}

// A vertex update makes the new vertices the mesh's rest pose (its node transform is the identity
// again); the device copy of the rest pose is refreshed on the next node transform.
void setRestPose(FhpScene_* s, uint32_t v0, uint32_t v1) {
    HostScene& h = s->h;
    std::copy(h.pos.begin() + 3 * size_t(v0), h.pos.begin() + 3 * size_t(v1), h.restPos.begin() + 3 * size_t(v0));
    std::copy(h.nrm.begin() + 3 * size_t(v0), h.nrm.begin() + 3 * size_t(v1), h.restNrm.begin() + 3 * size_t(v0));
    s->restValid = false;
}

extern "C" {

int fhp_get_version(void) { return 100; }

int fhp_session_create(int device, FhpSession* out) {
    if (!out) return fail(FHP_EINVAL, "null argument");
    int count = 0;
    HIP_OR_FAIL(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(FHP_EINVAL, "device ordinal out of range");
    HIP_OR_FAIL(hipSetDevice(device));
    hipStream_t st = nullptr;
    HIP_OR_FAIL(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));  // default priority
    *out = new FhpSession_{device, st};
    return FHP_OK;
}

int fhp_session_stream(FhpSession s, FhpStream* out) {
    if (!s || !out) return fail(FHP_EINVAL, "null argument");
    *out = FhpStream(s->stream);
    return FHP_OK;
}

int fhp_session_destroy(FhpSession s) {
    if (!s) return FHP_OK;
    const hipError_t e = hipStreamDestroy(s->stream);
    delete s;
    if (e != hipSuccess) return fail(FHP_EHIP, hipGetErrorString(e));
    return FHP_OK;
}

int fhp_pass_kernel_times(FhpPass p, double* ms, uint64_t* launches, int reset) {
    if (!p) return fail(FHP_EINVAL, "null argument");
    for (int i = 0; i < KK_COUNT; i++) {
        if (ms) ms[i] = p->timer.ms[i];
        if (launches) launches[i] = p->timer.launches[i];
        if (reset) { p->timer.ms[i] = 0.0; p->timer.launches[i] = 0; }
    }
    return FHP_OK;
}
const char* fhp_last_error(void) { return g_error.c_str(); }
const char* fhp_last_warning(void) { return g_warning.c_str(); }

// ------------------------------------------------------------------ scene
int fhp_scene_create(const FhpSceneDesc* d, FhpScene* out) {
    if (!d || !out) return fail(FHP_EINVAL, "null argument");
    std::unique_ptr<FhpScene_> s(new FhpScene_());
    std::string msg;
    if (int rc = sceneFromDesc(*d, s->h, msg)) return fail(rc, msg);
    int rc = buildAndUpload(s.get());
    if (rc) return rc;
    *out = s.release();
    return FHP_OK;
}

int fhp_scene_update_vertices(FhpScene s, uint32_t mesh, const float* pos, const float* nrm) {
    if (!s || !pos) return fail(FHP_EINVAL, "null argument");
    HostScene& h = s->h;
    if (mesh + 1 >= h.meshOff.size()) return fail(FHP_EINVAL, "mesh id out of range");
    const uint32_t v0 = h.meshOff[mesh], v1 = h.meshOff[mesh + 1];
    for (size_t i = 0; i < 3 * size_t(v1 - v0); i++)
        if (!std::isfinite(pos[i]) || (nrm && !std::isfinite(nrm[i]))) return fail(FHP_EINVAL, "non-finite vertex data");
    std::copy(pos, pos + 3 * size_t(v1 - v0), h.pos.begin() + 3 * size_t(v0));
    if (nrm) std::copy(nrm, nrm + 3 * size_t(v1 - v0), h.nrm.begin() + 3 * size_t(v0));
    setRestPose(s, v0, v1);
    return buildAndUpload(s);
}

int fhp_scene_update_vertices_device(FhpScene s, uint32_t mesh, const float* d_pos, const float* d_nrm, int rebuild,
                                     FhpStream stream) {
    if (!s || !d_pos) return fail(FHP_EINVAL, "null argument");
    HostScene& h = s->h;
    if (mesh + 1 >= h.meshOff.size()) return fail(FHP_EINVAL, "mesh id out of range");
    const uint32_t v0 = h.meshOff[mesh], v1 = h.meshOff[mesh + 1], nv = v1 - v0;
    hipStream_t st = hipStream_t(stream);
    // Host mirrors of the vertices (light collection, later rebuilds) and the finiteness check.
    std::vector<float> pos(3 * size_t(nv)), nrm(d_nrm ? 3 * size_t(nv) : 0);
    if (nv) {
        HIP_OR_FAIL(hipMemcpyAsync(pos.data(), d_pos, pos.size() * 4, hipMemcpyDeviceToHost, st));
        if (d_nrm) HIP_OR_FAIL(hipMemcpyAsync(nrm.data(), d_nrm, nrm.size() * 4, hipMemcpyDeviceToHost, st));
        HIP_OR_FAIL(hipStreamSynchronize(st));
    }
    for (float v : pos)
        if (!std::isfinite(v)) return fail(FHP_EINVAL, "non-finite vertex data");
    for (float v : nrm)
        if (!std::isfinite(v)) return fail(FHP_EINVAL, "non-finite vertex data");
    std::copy(pos.begin(), pos.end(), h.pos.begin() + 3 * size_t(v0));
    if (d_nrm) std::copy(nrm.begin(), nrm.end(), h.nrm.begin() + 3 * size_t(v0));
    setRestPose(s, v0, v1);
    // An emissive mesh also moves lights: rebuild the light collection and light BVH with the tree.
    bool emissive = false;
    for (uint32_t t = 0; t < h.nT && !emissive; t++) {
        const uint32_t i0 = h.idx[3 * size_t(t)];
        emissive = i0 >= v0 && i0 < v1 && h.triToLight[t] != kInvalid;
    }
    if (rebuild || emissive) return buildAndUpload(s);
    float ext = 0.f;
    for (float v : pos) ext = std::max(ext, std::fabs(v));
    h.bvhAbsPad = std::max(h.bvhAbsPad, ext * 1e-6f + 1e-30f);
    RefitArgs r{};
    r.vtx = s->vtx.as<float4>();
    r.triPos = s->triPos.as<float4>();
    r.bvh = s->bvh.as<float>();
    r.idx = s->idx.as<uint32_t>();
    r.pos = d_pos;
    r.nrm = d_nrm;
    r.v0 = v0;
    r.nVerts = nv;
    r.nTris = h.nT;
    r.levelNodes = s->levels.as<uint32_t>();
    r.levelOff = h.levelOff.data();
    r.nLevels = uint32_t(h.levelOff.size() - 1);
    r.absPad = h.bvhAbsPad;
    HIP_OR_FAIL(launchRefit(r, st));
    return FHP_OK;
}

int fhp_scene_update_node_transform(FhpScene s, uint32_t mesh, const float* h_matrix, FhpStream stream) {
    if (!s || !h_matrix) return fail(FHP_EINVAL, "null argument");
    HostScene& h = s->h;
    if (mesh + 1 >= h.meshOff.size()) return fail(FHP_EINVAL, "mesh id out of range");
    for (int i = 0; i < 12; i++)
        if (!std::isfinite(h_matrix[i])) return fail(FHP_EINVAL, "non-finite transform");
    // Normal matrix: inverse transpose of the 3x3 part (cofactors / determinant), in double.
    double a[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) a[r][c] = h_matrix[4 * r + c];
    const double c00 = a[1][1] * a[2][2] - a[1][2] * a[2][1], c01 = a[1][2] * a[2][0] - a[1][0] * a[2][2],
                 c02 = a[1][0] * a[2][1] - a[1][1] * a[2][0];
    const double det = a[0][0] * c00 + a[0][1] * c01 + a[0][2] * c02;
    if (!(det > 0.0))
        return fail(FHP_EINVAL, "node transform must have a positive determinant (mirroring would flip the mesh's winding)");
    const double cof[3][3] = {{c00, c01, c02},
                              {a[0][2] * a[2][1] - a[0][1] * a[2][2], a[0][0] * a[2][2] - a[0][2] * a[2][0],
                               a[0][1] * a[2][0] - a[0][0] * a[2][1]},
                              {a[0][1] * a[1][2] - a[0][2] * a[1][1], a[0][2] * a[1][0] - a[0][0] * a[1][2],
                               a[0][0] * a[1][1] - a[0][1] * a[1][0]}};
    NodeXformArgs x{};
    for (int i = 0; i < 12; i++) x.M[i] = h_matrix[i];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) x.N[3 * r + c] = float(cof[r][c] / det);  // (A^-1)^T = cof(A) / det
    const uint32_t v0 = h.meshOff[mesh], v1 = h.meshOff[mesh + 1], nv = v1 - v0;
    hipStream_t st = hipStream_t(stream);
    if (!s->restValid) {
        std::vector<float> rest(8 * size_t(h.nV));
        for (uint32_t i = 0; i < h.nV; i++) {
            float* q = &rest[8 * size_t(i)];
            q[0] = h.restPos[3 * i]; q[1] = h.restPos[3 * i + 1]; q[2] = h.restPos[3 * i + 2]; q[3] = h.restNrm[3 * i];
            q[4] = h.restNrm[3 * i + 1]; q[5] = h.restNrm[3 * i + 2]; q[6] = h.uv[2 * i]; q[7] = h.uv[2 * i + 1];
        }
        HIP_OR_FAIL(s->restVtx.upload(rest));
        s->restValid = true;
    }
    x.vtx = s->vtx.as<float4>();
    x.rest = s->restVtx.as<float4>();
    x.v0 = v0;
    x.nVerts = nv;
    HIP_OR_FAIL(launchNodeTransform(x, st));
    // Host mirror of the moved vertices (later rebuilds, vertex reads) and the box padding.
    std::vector<float> rec(8 * size_t(nv));
    if (nv) HIP_OR_FAIL(hipMemcpyAsync(rec.data(), s->vtx.as<float4>() + 2 * size_t(v0), rec.size() * 4, hipMemcpyDeviceToHost, st));
    HIP_OR_FAIL(hipStreamSynchronize(st));
    float ext = 0.f;
    for (uint32_t i = 0; i < nv; i++) {
        const float* q = &rec[8 * size_t(i)];
        float* p = &h.pos[3 * size_t(v0 + i)];
        float* n = &h.nrm[3 * size_t(v0 + i)];
        p[0] = q[0]; p[1] = q[1]; p[2] = q[2]; n[0] = q[3]; n[1] = q[4]; n[2] = q[5];
        for (int k = 0; k < 3; k++) {
            if (!std::isfinite(p[k])) return fail(FHP_EINVAL, "transformed vertices are not finite");
            ext = std::max(ext, std::fabs(p[k]));
        }
    }
    h.bvhAbsPad = std::max(h.bvhAbsPad, ext * 1e-6f + 1e-30f);
    RefitArgs r{};
    r.vtx = s->vtx.as<float4>();
    r.triPos = s->triPos.as<float4>();
    r.bvh = s->bvh.as<float>();
    r.idx = s->idx.as<uint32_t>();
    r.pos = nullptr;  // the vertex records already hold the new pose
    r.v0 = v0;
    r.nVerts = nv;
    r.nTris = h.nT;
    r.levelNodes = s->levels.as<uint32_t>();
    r.levelOff = h.levelOff.data();
    r.nLevels = uint32_t(h.levelOff.size() - 1);
    r.absPad = h.bvhAbsPad;
    HIP_OR_FAIL(launchRefit(r, st));
    if (h.nLights) {
        LightRefitArgs l{};
        l.lights = s->lights.as<float>();
        l.nLights = h.nLights;
        l.vtx = s->vtx.as<float4>();
        l.idx = s->idx.as<uint32_t>();
        l.lbvh = s->lbvh.as<uint32_t>();
        l.nNodes = h.nLbvh;
        l.lbvhTri = s->lbvhTri.as<uint32_t>();
        l.nodes = s->lrefit.as<uint32_t>();
        l.off = h.lrefitOff.data();
        l.depth = h.lrefitDepth;
        HIP_OR_FAIL(launchLightRefit(l, st));
        // Host mirrors of the light records and light BVH (fhp_scene_lights / fhp_scene_light_bvh).
        HIP_OR_FAIL(hipMemcpyAsync(h.lights.data(), s->lights.p, h.lights.size() * 4, hipMemcpyDeviceToHost, st));
        if (!h.lbvhPacked.empty())
            HIP_OR_FAIL(hipMemcpyAsync(h.lbvhPacked.data(), s->lbvh.p, h.lbvhPacked.size() * 4, hipMemcpyDeviceToHost, st));
        HIP_OR_FAIL(hipStreamSynchronize(st));
        decodeLightNodes(h);
    }
    return FHP_OK;
}

int fhp_bvh_build(FhpScene s, uint32_t max_leaf_triangles) {
    if (!s) return fail(FHP_EINVAL, "null argument");
    if (max_leaf_triangles > 8) return fail(FHP_EINVAL, "a BVH4 leaf holds at most 8 triangles");
    s->leafSize = max_leaf_triangles ? max_leaf_triangles : 4u;
    return buildAndUpload(s);
}

int fhp_scene_get_mesh_vertices(FhpScene s, uint32_t mesh, float* h_pos, float* h_nrm) {
    if (!s) return fail(FHP_EINVAL, "null argument");
    const HostScene& h = s->h;
    if (mesh + 1 >= h.meshOff.size()) return fail(FHP_EINVAL, "mesh id out of range");
    const uint32_t v0 = h.meshOff[mesh], v1 = h.meshOff[mesh + 1];
    if (h_pos) std::copy(h.pos.begin() + 3 * size_t(v0), h.pos.begin() + 3 * size_t(v1), h_pos);
    if (h_nrm) std::copy(h.nrm.begin() + 3 * size_t(v0), h.nrm.begin() + 3 * size_t(v1), h_nrm);
    return FHP_OK;
}

int fhp_scene_lights(FhpScene s, float* h_lights) {
    if (!s || !h_lights) return fail(FHP_EINVAL, "null argument");
    std::copy(s->h.lights.begin(), s->h.lights.end(), h_lights);
    return FHP_OK;
}

int fhp_scene_set_analytic_lights(FhpScene s, uint32_t n, const float* h_lights) {
    if (!s || (n && !h_lights)) return fail(FHP_EINVAL, "null argument");
    SceneDev& dv = s->dev;
    if (n == 0) {
        dv.alights = nullptr;
        dv.nALights = 0;
        return FHP_OK;
    }
    constexpr float kPi = 3.14159265358979323846f;
    std::vector<float> rec(16 * size_t(n), 0.f);
    for (uint32_t i = 0; i < n; i++) {
        const float* q = h_lights + 16 * size_t(i);
        for (int k = 0; k < 12; k++)
            if (!std::isfinite(q[k])) return fail(FHP_EINVAL, "non-finite analytic light " + std::to_string(i));
        if (!(q[0] == 0.f || q[0] == 1.f))
            return fail(FHP_EINVAL, "analytic light " + std::to_string(i) + ": type must be 0 (point) or 1 (directional)");
        const uint32_t type = q[0] == 0.f ? kLightPoint : kLightDirectional;
        // normalize() in float, one rounding per operation.
        const float xx = q[8] * q[8], yy = q[9] * q[9], zz = q[10] * q[10];
        float dd = xx + yy;
        dd = dd + zz;
        const float dl = std::sqrt(dd);
        if (!(dl > 0.f)) return fail(FHP_EINVAL, "analytic light " + std::to_string(i) + ": zero direction");
        // Light::setWorldDirection normalises; PointLight::setOpeningAngle / setPenumbraAngle clamp
        // (Light.cpp:174-223) and keep cos(openingAngle).
        const float opening = std::min(std::max(q[7], 0.f), kPi);
        const float penumbra = std::min(std::max(q[11], 0.f), opening);
        float* r = &rec[16 * size_t(i)];
        uint32_t t = type;
        std::memcpy(&r[0], &t, 4);
        r[1] = q[1]; r[2] = q[2]; r[3] = q[3];
        r[4] = q[4]; r[5] = q[5]; r[6] = q[6]; r[7] = opening;
        r[8] = q[8] / dl; r[9] = q[9] / dl; r[10] = q[10] / dl; r[11] = std::cos(opening);
        r[12] = penumbra;
    }
    if (int rc = waitForDevice()) return rc;
    HIP_OR_FAIL(s->alights.upload(rec));
    dv.alights = s->alights.as<float4>();
    dv.nALights = n;
    return FHP_OK;
}

int fhp_scene_set_env_map(FhpScene s, const float* h_texels, uint32_t W, uint32_t H, const float* h_params, uint32_t flags) {
    if (!s) return fail(FHP_EINVAL, "null argument");
    SceneDev& dv = s->dev;
    if (W == 0 || H == 0) {
        dv.envTex = nullptr;
        dv.envImp = nullptr;
        dv.envW = dv.envH = dv.envFlags = dv.envImpDim = 0;
        return FHP_OK;
    }
    if (!h_texels || !h_params) return fail(FHP_EINVAL, "null argument");
    if (flags & ~(kEnvLight | kEnvBackground)) return fail(FHP_EINVAL, "env map flags: 1 light, 2 background");
    const size_t n = 3 * size_t(W) * H;
    for (size_t i = 0; i < n; i++)
        if (!std::isfinite(h_texels[i])) return fail(FHP_EINVAL, "non-finite env map texel");
    if (int rc = waitForDevice()) return rc;
    HIP_OR_FAIL(s->envTex.alloc(n * 4));
    HIP_OR_FAIL(hipMemcpy(s->envTex.p, h_texels, n * 4, hipMemcpyHostToDevice));
    // h_params: intensity, tint rgb, rotation in degrees xyz. EnvMap::setRotation (EnvMap.cpp:62-72):
    // transform = matrixFromRotationXYZ (MatrixMath.h:649-679, angles negated); the world-to-local
    // rotation is its inverse, the transpose.
    const float kDeg = 3.14159265358979323846f / 180.f;
    const float c1 = std::cos(-h_params[4] * kDeg), c2 = std::cos(-h_params[5] * kDeg), c3 = std::cos(-h_params[6] * kDeg);
    const float s1 = std::sin(-h_params[4] * kDeg), s2 = std::sin(-h_params[5] * kDeg), s3 = std::sin(-h_params[6] * kDeg);
    const float m[3][3] = {{c2 * c3, c2 * s3, -s2},
                           {-c1 * s3 + s1 * s2 * c3, c1 * c3 + s1 * s2 * s3, s1 * c2},
                           {s1 * s3 + c1 * s2 * c3, -s1 * c3 + c1 * s2 * s3, c1 * c2}};
    dv.envInv0 = mk3(m[0][0], m[1][0], m[2][0]);
    dv.envInv1 = mk3(m[0][1], m[1][1], m[2][1]);
    dv.envInv2 = mk3(m[0][2], m[1][2], m[2][2]);
    dv.envScale = mk3(h_params[0] * h_params[1], h_params[0] * h_params[2], h_params[0] * h_params[3]);
    dv.envTex = s->envTex.as<float>();
    dv.envW = W;
    dv.envH = H;
    dv.envFlags = flags;
    // EnvMapSampler (EnvMapSampler.cpp:39-65): 512 x 512 importance map, 64 samples per texel.
    constexpr uint32_t kImpDim = 512, kImpSamplesPerAxis = 8;
    const size_t impN = (4 * size_t(kImpDim) * kImpDim - 1) / 3;
    HIP_OR_FAIL(s->envImp.alloc(impN * 4));
    HIP_OR_FAIL(launchEnvImportance(dv, s->envImp.as<float>(), kImpDim, kImpSamplesPerAxis, nullptr));
    float mean = 0.f;
    HIP_OR_FAIL(hipMemcpy(&mean, s->envImp.p, 4, hipMemcpyDeviceToHost));
    // A map whose luminance integrates to zero cannot be importance sampled (the descent would
    // divide by zero): env NEE is then off for this scene.
    const bool sampleable = mean > 0.f && std::isfinite(mean);
    dv.envImp = sampleable ? s->envImp.as<float>() : nullptr;
    dv.envImpDim = sampleable ? kImpDim : 0u;
    return FHP_OK;
}

int fhp_scene_get_env_importance(FhpScene s, uint32_t* dim, float* h_levels) {
    if (!s || !dim) return fail(FHP_EINVAL, "null argument");
    *dim = s->dev.envImpDim;
    if (h_levels && s->dev.envImp) {
        const size_t n = (4 * size_t(*dim) * *dim - 1) / 3;
        HIP_OR_FAIL(hipMemcpy(h_levels, s->dev.envImp, n * 4, hipMemcpyDeviceToHost));
    }
    return FHP_OK;
}

int fhp_env_sample(FhpScene s, const float* d_u, uint32_t n, float* d_out, FhpStream stream) {
    if (!s || (n && (!d_u || !d_out))) return fail(FHP_EINVAL, "null argument");
    if (!s->dev.envImp) return fail(FHP_ESTATE, "the scene has no env map that can be importance sampled");
    if (n) HIP_OR_FAIL(launchEnvSampleBatch(s->dev, d_u, n, d_out, hipStream_t(stream)));
    return FHP_OK;
}

int fhp_env_eval_pdf(FhpScene s, const float* d_dirs, uint32_t n, float* d_pdf, FhpStream stream) {
    if (!s || (n && (!d_dirs || !d_pdf))) return fail(FHP_EINVAL, "null argument");
    if (!s->dev.envImp) return fail(FHP_ESTATE, "the scene has no env map that can be importance sampled");
    if (n) HIP_OR_FAIL(launchEnvPdfBatch(s->dev, d_dirs, n, d_pdf, hipStream_t(stream)));
    return FHP_OK;
}

int fhp_scene_destroy(FhpScene s) {
    if (!s) return FHP_OK;
    const int rc = waitForDevice();
    delete s;
    return rc;
}

int fhp_scene_trace_errors(FhpScene s, uint32_t* bits, int reset, FhpStream stream) {
    if (!s || !bits) return fail(FHP_EINVAL, "null argument");
    hipStream_t st = hipStream_t(stream);
    uint32_t v = 0;
    HIP_OR_FAIL(hipMemcpyAsync(&v, s->err.p, 4, hipMemcpyDeviceToHost, st));
    HIP_OR_FAIL(hipStreamSynchronize(st));
    if (reset && v) HIP_OR_FAIL(launchFill(s->err.as<float>(), 1, 0.f, st));
    *bits = v;
    return FHP_OK;
}

int fhp_scene_set_camera(FhpScene s, const float* h_camera) {
    if (!s || !h_camera) return fail(FHP_EINVAL, "null argument");
    for (int i = 0; i < 14; i++)
        if (!std::isfinite(h_camera[i])) return fail(FHP_EINVAL, "non-finite camera parameter");
    HostScene& h = s->h;
    std::memcpy(h.cam, h_camera, sizeof(h.cam));
    computeCamera(h);
    SceneDev& dv = s->dev;
    dv.cam.pos = mk3(h.cam[0], h.cam[1], h.cam[2]);
    dv.cam.U = mk3(h.camU[0], h.camU[1], h.camU[2]);
    dv.cam.V = mk3(h.camV[0], h.camV[1], h.camV[2]);
    dv.cam.W = mk3(h.camW[0], h.camW[1], h.camW[2]);
    dv.cam.nearZ = h.cam[12];
    dv.cam.farZ = h.cam[13];
    dv.viewProjZ = make_float4(h.viewProj[8], h.viewProj[9], h.viewProj[10], h.viewProj[11]);
    dv.viewProjW = make_float4(h.viewProj[12], h.viewProj[13], h.viewProj[14], h.viewProj[15]);
    return FHP_OK;
}

int fhp_scene_info(FhpScene s, FhpSceneInfo* o) {
    if (!s || !o) return fail(FHP_EINVAL, "null argument");
    const HostScene& h = s->h;
    o->num_triangles = h.nT; o->num_vertices = h.nV; o->num_materials = h.nM; o->num_textures = h.nTex;
    o->bvh_nodes = h.bvhNodes; o->bvh_leaves = h.bvhLeaves; o->bvh_depth = h.bvhDepth; o->stack_need = h.stackNeed;
    o->num_lights = h.nLights; o->num_active_lights = uint32_t(h.activeLights.size()); o->light_bvh_nodes = h.nLbvh;
    o->num_meshes = uint32_t(h.meshOff.size() - 1);
    o->hit_info_compressed = h.hitCompressed;
    o->hit_info_type_bits = h.hitTypeBits;
    o->hit_info_instance_bits = h.hitInstBits;
    o->hit_info_primitive_bits = h.hitPrimBits;
    o->meshes_contiguous = h.meshTriContiguous;
    o->bvh_sah_cost = h.bvhSah;
    o->device_bytes = s->deviceBytes;
    for (int i = 0; i < 3; i++) {
        o->camera_basis[i] = h.camU[i]; o->camera_basis[3 + i] = h.camV[i]; o->camera_basis[6 + i] = h.camW[i];
    }
    return FHP_OK;
}

int fhp_scene_get_materials(FhpScene s, float* mats) {
    if (!s || !mats) return fail(FHP_EINVAL, "null argument");
    std::copy(s->h.mats.begin(), s->h.mats.end(), mats);
    return FHP_OK;
}

int fhp_scene_set_materials(FhpScene s, const float* mats) {
    if (!s || !mats) return fail(FHP_EINVAL, "null argument");
    for (uint32_t m = 0; m < s->h.nM; m++) {
        int tex = int(mats[size_t(m) * M_STRIDE + M_TEX]);
        if (tex >= int(s->h.nTex)) return fail(FHP_EINVAL, "material texture index out of range");
        if (int(mats[size_t(m) * M_STRIDE + M_EMIT_TEX]) - 1 >= int(s->h.nTex))
            return fail(FHP_EINVAL, "material emissive texture index out of range");
    }
    if (int rc = waitForDevice()) return rc;
    s->h.mats.assign(mats, mats + size_t(s->h.nM) * M_STRIDE);
    HIP_OR_FAIL(hipMemcpy(s->mats.p, mats, size_t(s->h.nM) * M_STRIDE * 4, hipMemcpyHostToDevice));
    s->dev.hasTransmission = hasTransmission(s->h) ? 1u : 0u;
    // Light samples evaluate the emitters' current materials (EmissiveLightSamplerHelpers.slang:79).
    if (s->h.nLights) {
        refreshLightEmission(s->h);
        HIP_OR_FAIL(hipMemcpy(s->lights.p, s->h.lights.data(), s->h.lights.size() * 4, hipMemcpyHostToDevice));
    }
    return FHP_OK;
}

// StandardMaterialParamLayout.slang:32-41: serialized offset, size, record field, deserialize
// clamp range (MaterialParamLayout.slangh:40-45, defaults [1e-4, 1 - 1e-4]) and whether the
// material stores the value in float16 (BasicMaterialData.slang:50-66, MaterialData.slang:183).
struct ParamEntry { int offset, size, field; float lo, hi; bool half; };
constexpr float kParamEps = 1e-4f;
constexpr ParamEntry kParamLayout[] = {
    {0, 3, M_BASE, kParamEps, 1.f - kParamEps, true},        // baseColor
    {3, 1, M_METAL, kParamEps, 1.f - kParamEps, true},       // metallic
    {4, 1, M_ROUGH, 0.05f, 1.f - kParamEps, true},           // roughness
    {5, 1, M_IOR, 1e-3f, 1e3f, true},                        // ior
    {6, 3, M_TRANS, kParamEps, 1.f - kParamEps, true},       // transmissionColor
    {9, 1, M_DIFF_TRANS, 0.f, 1.f - kParamEps, true},        // diffuseTransmission
    {10, 1, M_SPEC_TRANS, 0.f, 1.f - kParamEps, true},       // specularTransmission
    {11, 3, M_EMIT, kParamEps, 1.f - kParamEps, false},      // emissiveColor
    {14, 1, M_EMIT_FACTOR, 1e-8f, 1e8f, false},              // emissiveFactor
};
constexpr int kParamCount = 20;  // SerializedMaterialParams::kParamCount

int fhp_scene_get_material_params(FhpScene s, uint32_t n, const uint32_t* ids, float* params) {
    if (!s || (n && (!ids || !params))) return fail(FHP_EINVAL, "null argument");
    for (uint32_t i = 0; i < n; i++) {
        if (ids[i] >= s->h.nM) return fail(FHP_EINVAL, "material id out of range");
        const float* r = &s->h.mats[size_t(ids[i]) * M_STRIDE];
        if (int(r[M_TYPE]) != MAT_STANDARD)
            return fail(FHP_EINVAL, "material " + std::to_string(ids[i]) + " does not support serializing parameters");
        float* o = params + size_t(i) * kParamCount;
        std::fill(o, o + kParamCount, 0.f);
        for (const ParamEntry& e : kParamLayout)
            for (int k = 0; k < e.size; k++) o[e.offset + k] = r[e.field + k];
    }
    return FHP_OK;
}

int fhp_scene_set_material_params(FhpScene s, uint32_t n, const uint32_t* ids, const float* params) {
    if (!s || (n && (!ids || !params))) return fail(FHP_EINVAL, "null argument");
    std::vector<float> mats = s->h.mats;
    for (uint32_t i = 0; i < n; i++) {
        if (ids[i] >= s->h.nM) return fail(FHP_EINVAL, "material id out of range");
        float* r = &mats[size_t(ids[i]) * M_STRIDE];
        if (int(r[M_TYPE]) != MAT_STANDARD)
            return fail(FHP_EINVAL, "material " + std::to_string(ids[i]) + " does not support deserializing parameters");
        const float* q = params + size_t(i) * kParamCount;
        for (const ParamEntry& e : kParamLayout)
            for (int k = 0; k < e.size; k++) {
                const float v = q[e.offset + k];
                if (std::isnan(v)) return fail(FHP_EINVAL, "NaN material parameter");
                const float c = std::min(std::max(v, e.lo), e.hi);
                r[e.field + k] = e.half ? float(_Float16(c)) : c;
            }
    }
    return fhp_scene_set_materials(s, mats.data());
}

int fhp_scene_set_texels(FhpScene s, const float* d_texels, uint64_t n, FhpStream stream) {
    if (!s || !d_texels) return fail(FHP_EINVAL, "null argument");
    if (n != s->h.texels.size()) return fail(FHP_EINVAL, "texel count mismatch");
    HIP_OR_FAIL(hipMemcpyAsync(s->texels.p, d_texels, n * 4, hipMemcpyDeviceToDevice, hipStream_t(stream)));
    return FHP_OK;
}

int fhp_scene_freeze_sampling(FhpScene s, int enable, FhpStream stream) {
    if (!s) return fail(FHP_EINVAL, "null argument");
    hipStream_t st = hipStream_t(stream);
    if (!enable) {
        s->dev.matsS = s->dev.mats;
        s->dev.texelsS = s->dev.texels;
        return FHP_OK;
    }
    if (s->matsS.bytes < s->mats.bytes || s->texelsS.bytes < s->texels.bytes)
        if (int rc = waitForDevice()) return rc;  // the frozen copies are reallocated
    HIP_OR_FAIL(s->matsS.alloc(s->mats.bytes));
    HIP_OR_FAIL(hipMemcpyAsync(s->matsS.p, s->mats.p, s->mats.bytes, hipMemcpyDeviceToDevice, st));
    HIP_OR_FAIL(s->texelsS.alloc(s->texels.bytes));
    HIP_OR_FAIL(hipMemcpyAsync(s->texelsS.p, s->texels.p, s->texels.bytes, hipMemcpyDeviceToDevice, st));
    s->dev.matsS = s->matsS.as<float>();
    s->dev.texelsS = s->texelsS.as<float>();
    return FHP_OK;
}

int fhp_scene_light_bvh(FhpScene s, float* nodes, uint32_t* tri, uint32_t* masks) {
    if (!s) return fail(FHP_EINVAL, "null argument");
    if (nodes) std::copy(s->h.lbvh.begin(), s->h.lbvh.end(), nodes);
    if (tri) std::copy(s->h.lbvhTri.begin(), s->h.lbvhTri.end(), tri);
    if (masks) std::copy(s->h.lbvhMask.begin(), s->h.lbvhMask.end(), masks);
    return FHP_OK;
}

// ------------------------------------------------------------------ ray queries
int fhp_trace_closest(FhpScene s, const float* d_rays, uint32_t n, uint32_t* d_hits, FhpStream stream) {
    if (!s || (n && (!d_rays || !d_hits))) return fail(FHP_EINVAL, "null argument");
    HIP_OR_FAIL(launchTraceRays(s->dev, reinterpret_cast<const float4*>(d_rays), n, reinterpret_cast<uint4*>(d_hits), nullptr,
                                false, s->ts, hipStream_t(stream)));
    return FHP_OK;
}
int fhp_trace_any(FhpScene s, const float* d_rays, uint32_t n, uint32_t* d_occ, FhpStream stream) {
    if (!s || (n && (!d_rays || !d_occ))) return fail(FHP_EINVAL, "null argument");
    HIP_OR_FAIL(launchTraceRays(s->dev, reinterpret_cast<const float4*>(d_rays), n, nullptr, d_occ, true, s->ts,
                                hipStream_t(stream)));
    return FHP_OK;
}

// ------------------------------------------------------------------ passes
int fhp_pass_create(const char* kind, int np, const char* const* keys, const double* values, FhpPass* out) {
    if (!kind || !out || (np > 0 && (!keys || !values))) return fail(FHP_EINVAL, "null argument");
    std::string k(kind);
    const std::map<std::string, double>* defs;
    if (k == "MinimalPathTracer") defs = &mptDefaults();
    else if (k == "WARDiffPathTracer") defs = &warDefaults();
    else if (k == "AccumulatePass") defs = &accDefaults();
    else if (k == "VBufferRT") defs = &vbufDefaults();
    else return fail(FHP_EINVAL, "unknown pass kind '" + k + "'");
    std::unique_ptr<FhpPass_> p(new FhpPass_());
    p->kind = k;
    p->props = *defs;
    g_warning.clear();
    const bool acc = k == "AccumulatePass";
    bool hasEnabled = false;
    int legacyEnable = -1;
    for (int i = 0; i < np; i++) {
        if (!keys[i]) return fail(FHP_EINVAL, "null property key");
        std::string key(keys[i]);
        if (acc && key == "enableAccumulation") {
            // Deprecated alias of 'enabled' (AccumulatePass.cpp:87-92).
            g_warning = "'enableAccumulation' is deprecated. Use 'enabled' instead.";
            legacyEnable = i;
            continue;
        }
        if (!defs->count(key)) {
            g_warning = "Unknown property '" + key + "' in " + k + " properties.";
            continue;
        }
        if (acc) {
            const std::string msg = accCheck(key, values[i]);
            if (!msg.empty()) return fail(FHP_EINVAL, msg);
            hasEnabled |= key == "enabled";
        }
        if (k == "VBufferRT") {
            const std::string msg = vbufCheck(key, values[i], p->props);
            if (!msg.empty()) return fail(FHP_EINVAL, msg);
        }
        p->props[key] = values[i];
        if (key == "fixedSeed") p->props["useFixedSeed"] = 1;
    }
    if (acc) {
        if (legacyEnable >= 0 && !hasEnabled) p->props["enabled"] = values[legacyEnable];
        *out = p.release();
        return FHP_OK;
    }
    if (k == "VBufferRT")
        p->props["sampleCount"] = p->jitter.create(int(p->get("samplePattern")), uint32_t(p->get("sampleCount")));
    if (k == "WARDiffPathTracer") {
        const double dm = p->get("diffMode"), dv = p->get("diffVarName");
        if (!(dm == 0 || dm == 1 || dm == 2 || dm == 3))
            return fail(FHP_EINVAL, "diffMode: 0 Primal, 1 BackwardDiff, 2 ForwardDiffDebug, 3 BackwardDiffDebug");
        if (!(dv == 0 || dv == 1 || dv == 2))
            return fail(FHP_EINVAL, "diffVarName: 0 none, 1 CBOX_BUNNY_MATERIAL, 2 CBOX_BUNNY_TRANSLATION");
        applyDiffVarName(p.get());
    }
    if (p->get("maxBounces") < 0 || p->get("maxBounces") > kMaxBounceSlots - 2)
        return fail(FHP_EINVAL, "maxBounces out of range [0, 28]");
    if (!validLaunch(p->get("traceBlockSize"), p->get("traceStackDepth"), p->get("traceBlocksPerCU")))
        return fail(FHP_EINVAL, "traversal launch: traceBlockSize in {64,128,256}, traceStackDepth in {16,24,32}, "
                                "traceBlocksPerCU >= 1 with traceBlocksPerCU*traceBlockSize <= 2048");
    *out = p.release();
    return FHP_OK;
}
int fhp_pass_destroy(FhpPass p) {
    delete p;
    return FHP_OK;
}
int fhp_pass_set_property(FhpPass p, const char* key, double v) {
    if (!p || !key) return fail(FHP_EINVAL, "null argument");
    if (!p->props.count(key)) {
        g_warning = std::string("Unknown property '") + key + "' in " + p->kind + " properties.";
        return FHP_OK;
    }
    const std::string k(key);
    if (p->kind == "AccumulatePass") {
        const std::string msg = accCheck(k, v);
        if (!msg.empty()) return fail(FHP_EINVAL, msg);
        // Toggling accumulation or changing the precision mode restarts it
        // (AccumulatePass::setEnabled, :366-373; the new mode's sums are created cleared, :380-420).
        if ((k == "enabled" && (v != 0) != (p->get("enabled") != 0)) || (k == "precisionMode" && v != p->get("precisionMode")))
            p->accCount = 0;
        p->props[k] = v;
        return FHP_OK;
    }
    if (p->kind == "VBufferRT") {
        const std::string msg = vbufCheck(k, v, p->props);
        if (!msg.empty()) return fail(FHP_EINVAL, msg);
        if (k == "samplePattern" || k == "sampleCount") {
            // A new pattern generator (GBufferBase::updateSamplePattern, GBufferBase.cpp:231-236).
            p->props[k] = v;
            p->props["sampleCount"] = p->jitter.create(int(p->get("samplePattern")), uint32_t(p->get("sampleCount")));
            return FHP_OK;
        }
    }
    if (k == "maxBounces" && (v < 0 || v > kMaxBounceSlots - 2)) return fail(FHP_EINVAL, "maxBounces out of range");
    if (p->kind == "WARDiffPathTracer" && k == "diffMode" && !(v == 0 || v == 1 || v == 2 || v == 3))
        return fail(FHP_EINVAL, "diffMode: 0 Primal, 1 BackwardDiff, 2 ForwardDiffDebug, 3 BackwardDiffDebug");
    if (p->kind == "WARDiffPathTracer" && k == "diffVarName") {
        if (!(v == 0 || v == 1 || v == 2))
            return fail(FHP_EINVAL, "diffVarName: 0 none, 1 CBOX_BUNNY_MATERIAL, 2 CBOX_BUNNY_TRANSLATION");
        p->props[k] = v;
        applyDiffVarName(p);
        return FHP_OK;
    }
    if (k == "tailPaths" && !(v >= 0 && v <= 16777216 && v == double(uint32_t(v))))
        return fail(FHP_EINVAL, "tailPaths: an integer in [0, 2^24]");
    if (k == "tailLanes" && !(v >= 1 && v <= 32 && v == double(uint32_t(v))))
        return fail(FHP_EINVAL, "tailLanes: an integer in [1, 32]");
    if (k == "traceBlockSize" || k == "traceStackDepth" || k == "traceBlocksPerCU") {
        const double b = k == "traceBlockSize" ? v : p->get("traceBlockSize");
        const double l = k == "traceStackDepth" ? v : p->get("traceStackDepth");
        const double c = k == "traceBlocksPerCU" ? v : p->get("traceBlocksPerCU");
        if (!validLaunch(b, l, c))
            return fail(FHP_EINVAL, "traversal launch: traceBlockSize in {64,128,256}, traceStackDepth in {16,24,32}, "
                                    "traceBlocksPerCU >= 1 with traceBlocksPerCU*traceBlockSize <= 2048");
    }
    p->props[key] = v;
    if (std::string(key) == "fixedSeed") p->props["useFixedSeed"] = 1;
    return FHP_OK;
}
int fhp_pass_get_property(FhpPass p, const char* key, double* v) {
    if (!p || !key || !v) return fail(FHP_EINVAL, "null argument");
    auto it = p->props.find(key);
    if (it == p->props.end()) return fail(FHP_EINVAL, std::string("no property '") + key + "'");
    *v = it->second;
    return FHP_OK;
}

int fhp_pass_execute(FhpPass p, FhpScene s, uint32_t W, uint32_t H, uint32_t spp, uint32_t frame, float* d_color,
                     uint64_t* h_stats, FhpStream stream) {
    if (!p || !d_color) return fail(FHP_EINVAL, "null argument");
    if (p->kind == "AccumulatePass") return fail(FHP_EINVAL, "AccumulatePass runs through fhp_accumulate");
    if (p->kind == "VBufferRT") return fail(FHP_EINVAL, "VBufferRT runs through fhp_vbuffer_execute");
    if (W == 0 || H == 0 || spp == 0) return fail(FHP_EINVAL, "empty frame");
    if (W > 65536 || H > 65536) return fail(FHP_EINVAL, "frame dimension above 65536");
    hipStream_t st = hipStream_t(stream);
    if (!s) {
        // No scene: clear the output (MinimalPathTracer.cpp:121-130).
        HIP_OR_FAIL(launchFill(d_color, uint64_t(W) * H * 4, 0.f, st));
        return FHP_OK;
    }
    int rc = ensurePassBuffers(p, W, H, spp);
    if (rc) return rc;
    const bool count = p->get("countTraversal") != 0;
    unsigned long long* dstats = p->stats.as<unsigned long long>();
    if (h_stats) HIP_OR_FAIL(launchFill(reinterpret_cast<float*>(dstats), 16, 0.f, st));
    TraceScratch ts;
    if ((rc = passTrace(s, p, ts))) return rc;
    if (p->kind == "MinimalPathTracer") {
        MptArgs a;
        if ((rc = mptFrameArgs(p, s, W, H, spp, frame, a))) return rc;
        a.thp = p->thp.as<float4>(); a.prevN = p->prevN.as<float4>(); a.rng = p->rng.as<uint4>();
        a.Lp = p->Lp.as<float4>(); a.Ls = p->Ls.as<float4>();
        a.rayQ[0] = p->rayQ0.as<float4>(); a.rayQ[1] = p->rayQ1.as<float4>();
        a.pathQ[0] = p->pathQ0.as<uint32_t>(); a.pathQ[1] = p->pathQ1.as<uint32_t>();
        a.hitQ = p->hitQ.as<uint4>(); a.hitIdx = p->hitIdx.as<uint32_t>(); a.shQ = p->shQ.as<float4>(); a.pixHit = p->pixHit.as<uint4>();
        a.counters = p->counters.as<uint32_t>();
        if ((rc = prepareTimer(p))) return rc;
        const bool graph = p->get("useGraph") != 0 && !count && !p->timer.on;
        if (!graph) {
            HIP_OR_FAIL(launchMptFrame(s->dev, a, ts, reinterpret_cast<float4*>(d_color), count,
                                       h_stats ? dstats : nullptr, st, &p->timer));
        } else {
            // Key: scene and argument bytes with the frame number zeroed (the kernels read it from
            // device memory, written by the prologue).
            MptArgs ka;
            std::memcpy(&ka, &a, sizeof(MptArgs));
            ka.frame = 0;
            std::vector<uint8_t> key(sizeof(SceneDev) + sizeof(MptArgs) + sizeof(TraceScratch));
            std::memcpy(key.data(), &s->dev, sizeof(SceneDev));
            std::memcpy(key.data() + sizeof(SceneDev), &ka, sizeof(MptArgs));
            std::memcpy(key.data() + sizeof(SceneDev) + sizeof(MptArgs), &ts, sizeof(TraceScratch));
            if (key != p->graphKey) {
                p->dropGraph();
                if (!p->capStream) HIP_OR_FAIL(hipStreamCreateWithFlags(&p->capStream, hipStreamNonBlocking));
                hipGraph_t g = nullptr;
                HIP_OR_FAIL(hipStreamBeginCapture(p->capStream, hipStreamCaptureModeThreadLocal));
                hipError_t ce = launchMptBounces(s->dev, a, ts, false, nullptr, p->capStream, nullptr);
                hipError_t ee = hipStreamEndCapture(p->capStream, &g);
                HIP_OR_FAIL(ce);
                HIP_OR_FAIL(ee);
                hipError_t ie = hipGraphInstantiate(&p->graph, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                HIP_OR_FAIL(ie);
                p->graphKey = key;
            }
            HIP_OR_FAIL(launchMptPrologue(a, st));
            HIP_OR_FAIL(hipGraphLaunch(p->graph, st));
            HIP_OR_FAIL(launchMptEpilogue(s->dev, a, reinterpret_cast<float4*>(d_color), h_stats ? dstats : nullptr, st));
        }
        if ((rc = collectTimer(p, st))) return rc;
    } else {
        WarArgs a{};
        a.W = W; a.H = H; a.spp = spp;
        a.frame = p->get("useFixedSeed") != 0 ? uint32_t(p->get("fixedSeed")) : frame * spp;
        a.maxBounces = uint32_t(p->get("maxBounces"));
        a.useBSDFSampling = p->get("useBSDFSampling") != 0;
        a.useNEE = p->get("useNEE") != 0;
        a.useMIS = p->get("useMIS") != 0;
        a.useAntithetic = p->get("useAntitheticSampling") != 0;
        a.out = reinterpret_cast<float4*>(d_color);
        a.tmp = p->Lp.as<float4>();
        a.rays = h_stats ? dstats : nullptr;
        const int mode = int(p->get("diffMode"));
        const uint64_t npix = uint64_t(W) * H;
        if ((rc = prepareTimer(p))) return rc;
        if (mode == kDiffBackwardDebug) {
            // BackwardDiffDebug renders no colour (the reference clears its outputs); its gradients
            // come from fhp_pass_backward.
            HIP_OR_FAIL(launchFill(d_color, npix * 4, 0.f, st));
            if (p->outDColor) HIP_OR_FAIL(launchFill(reinterpret_cast<float*>(p->outDColor), npix * 4, 0.f, st));
        } else if (mode == kDiffForwardDebug && p->dbgVar == kDiffVarMaterial) {
            a.dbgMat = p->dbgId;
            a.dbgOffset = p->dbgOffset;
            std::memcpy(a.dbgGrad, p->dbgGrad, sizeof(a.dbgGrad));
            a.dbgThreshold = float(p->get("assertThreshold"));
            a.dColor = p->outDColor;
            if ((rc = prepareHitLog(p, s, ts, a))) return rc;
            HIP_OR_FAIL(launchWarDebugImage(s->dev, a, ts, st, &p->timer));
        } else {
            if (mode == kDiffForwardDebug && p->dbgVar == kDiffVarGeometryTranslation)
                return fail(FHP_EINVAL, "ForwardDiffDebug images of a GeometryTranslation variable are not supported");
            HIP_OR_FAIL(launchWarPrimal(s->dev, a, ts, st, &p->timer));
            // ForwardDiffDebug without a debug variable: every derivative is zero (written, as the
            // reference writes (0, 0, 0, 1) for a finite zero derivative).
            if (mode == kDiffForwardDebug && p->outDColor) HIP_OR_FAIL(launchFillRGBA(p->outDColor, npix, 0.f, 0.f, 0.f, 1.f, st));
        }
        if ((rc = collectTimer(p, st))) return rc;
    }
    if (h_stats) {
        unsigned long long hs[8];
        HIP_OR_FAIL(hipMemcpyAsync(hs, dstats, 64, hipMemcpyDeviceToHost, st));
        HIP_OR_FAIL(hipStreamSynchronize(st));
        for (int i = 0; i < 8; i++) h_stats[i] = hs[i];
    }
    return FHP_OK;
}

// ------------------------------------------------------------------ allocation-free forward frame
// fhp_pass_execute_noalloc keeps a MinimalPathTracer frame's wavefront state (path state, ray,
// hit and visibility queues, primary hits, queue counters) in this file-scope device array
// instead of pass-owned buffers. Each device holds its own copy (kNoAllocScratchBytes, about
// 124 MB) from module load until the process ends.
constexpr uint64_t kNoAllocMaxPaths = 1ull << 19, kNoAllocMaxPixels = 1ull << 19;
constexpr uint64_t align256(uint64_t x) { return (x + 255u) & ~uint64_t(255u); }
// thp, prevN, rng, Lp, Ls (16 B each), two ray queues (32 B), two path queues (4 B), hits (16 B),
// visibility records (48 B), hit indices (4 B) per path; primary hits (16 B) per pixel; counters.
constexpr uint64_t kNoAllocScratchBytes =
    5 * align256(16 * kNoAllocMaxPaths) + 2 * align256(32 * kNoAllocMaxPaths) + 2 * align256(4 * kNoAllocMaxPaths) +
    align256(16 * kNoAllocMaxPaths) + align256(48 * kNoAllocMaxPaths) + align256(4 * kNoAllocMaxPaths) +
    align256(16 * kNoAllocMaxPixels) + align256(4 * kCounterWords);
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_scratchMutex;
constexpr int kMaxDevices = 64;
uint8_t* g_scratchPtr[kMaxDevices] = {};

// Device address of this device's scratch array, looked up once per device (with the device
// current) on the entry point's first call there. In host code the array's name is not a device
// address; only this pointer is passed to kernels.
int scratchAddress(uint8_t** out) {
    int dev = 0;
    HIP_OR_FAIL(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(FHP_EINVAL, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (!g_scratchPtr[dev]) {
        void* ptr = nullptr;
        HIP_OR_FAIL(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return fail(FHP_EHIP, "scratch array has no device address");
        g_scratchPtr[dev] = static_cast<uint8_t*>(ptr);
    }
    *out = g_scratchPtr[dev];
    return FHP_OK;
}
}  // namespace

int fhp_pass_execute_noalloc(FhpPass p, FhpScene s, uint32_t W, uint32_t H, uint32_t spp, uint32_t frame, float* d_color,
                             FhpStream stream) {
    if (!p || !s || !d_color) return fail(FHP_EINVAL, "null argument");
    if (p->kind != "MinimalPathTracer") return fail(FHP_EINVAL, "fhp_pass_execute_noalloc runs MinimalPathTracer passes");
    if (W == 0 || H == 0 || spp == 0) return fail(FHP_EINVAL, "empty frame");
    const uint64_t nPix = uint64_t(W) * H, nPaths = nPix * spp;
    if (nPaths > kNoAllocMaxPaths || nPix > kNoAllocMaxPixels)
        return fail(FHP_EINVAL, "frame too large for fhp_pass_execute_noalloc (at most " + std::to_string(kNoAllocMaxPaths) +
                                    " paths and " + std::to_string(kNoAllocMaxPixels) + " pixels)");
    if (p->get("countTraversal") != 0) return fail(FHP_EINVAL, "fhp_pass_execute_noalloc does not count traversal");
    MptArgs a;
    int rc = mptFrameArgs(p, s, W, H, spp, frame, a);
    if (rc) return rc;
    TraceScratch ts;
    if ((rc = passTrace(s, p, ts))) return rc;
    uint8_t* base = nullptr;
    if ((rc = scratchAddress(&base))) return rc;
    uint64_t off = 0;
    auto take = [&](uint64_t bytes) {
        uint8_t* q = base + off;
        off += align256(bytes);
        return q;
    };
    const uint64_t n = kNoAllocMaxPaths;
    a.thp = reinterpret_cast<float4*>(take(16 * n));
    a.prevN = reinterpret_cast<float4*>(take(16 * n));
    a.rng = reinterpret_cast<uint4*>(take(16 * n));
    a.Lp = reinterpret_cast<float4*>(take(16 * n));
    a.Ls = reinterpret_cast<float4*>(take(16 * n));
    a.rayQ[0] = reinterpret_cast<float4*>(take(32 * n));
    a.rayQ[1] = reinterpret_cast<float4*>(take(32 * n));
    a.pathQ[0] = reinterpret_cast<uint32_t*>(take(4 * n));
    a.pathQ[1] = reinterpret_cast<uint32_t*>(take(4 * n));
    a.hitQ = reinterpret_cast<uint4*>(take(16 * n));
    a.shQ = reinterpret_cast<float4*>(take(48 * n));
    a.hitIdx = reinterpret_cast<uint32_t*>(take(4 * n));
    a.pixHit = reinterpret_cast<uint4*>(take(16 * kNoAllocMaxPixels));
    a.counters = reinterpret_cast<uint32_t*>(take(4 * kCounterWords));
    if (off > kNoAllocScratchBytes) return fail(FHP_EHIP, "scratch layout exceeds the scratch array");
    // The frame clears its own counters and writes every path slot before reading it, so the
    // previous call's contents need no clearing.
    HIP_OR_FAIL(launchMptFrame(s->dev, a, ts, reinterpret_cast<float4*>(d_color), false, nullptr, hipStream_t(stream),
                               nullptr));
    return FHP_OK;
}

static int warExecuteBatch(FhpPass p, uint32_t nv, const FhpView* views, uint64_t* h_stats, FhpStream stream);

int fhp_pass_execute_batch(FhpPass p, uint32_t nv, const FhpView* views, uint64_t* h_stats, FhpStream stream) {
    if (!p || (nv && !views)) return fail(FHP_EINVAL, "null argument");
    if (nv == 0) return FHP_OK;
    if (p->kind == "WARDiffPathTracer") return warExecuteBatch(p, nv, views, h_stats, stream);
    if (p->kind != "MinimalPathTracer")
        return fail(FHP_EINVAL, "batched execution is available for MinimalPathTracer and WARDiffPathTracer");
    if (nv > 65535) return fail(FHP_EINVAL, "at most 65535 views per batch");
    hipStream_t st = hipStream_t(stream);
    std::vector<ViewDev> vt(nv);
    std::vector<FhpScene> scenes;
    uint64_t nPix = 0, nPaths = 0;
    FhpScene deepest = nullptr;
    for (uint32_t i = 0; i < nv; i++) {
        const FhpView& v = views[i];
        if (!v.scene || !v.d_color) return fail(FHP_EINVAL, "view " + std::to_string(i) + ": null scene or output");
        if (v.width == 0 || v.height == 0 || v.spp == 0 || v.width > 65536 || v.height > 65536)
            return fail(FHP_EINVAL, "view " + std::to_string(i) + ": bad frame shape");
        uint32_t si = 0;
        while (si < scenes.size() && scenes[si] != v.scene) si++;
        if (si == scenes.size()) scenes.push_back(v.scene);
        if (!deepest || v.scene->stackNeed > deepest->stackNeed) deepest = v.scene;
        ViewDev& d = vt[i];
        d.cam = v.scene->dev.cam;
        if (v.h_camera) {
            HostScene c;
            std::memcpy(c.cam, v.h_camera, sizeof(c.cam));
            computeCamera(c);
            d.cam.pos = mk3(c.cam[0], c.cam[1], c.cam[2]);
            d.cam.U = mk3(c.camU[0], c.camU[1], c.camU[2]);
            d.cam.V = mk3(c.camV[0], c.camV[1], c.camV[2]);
            d.cam.W = mk3(c.camW[0], c.camW[1], c.camW[2]);
            d.cam.nearZ = c.cam[12];
            d.cam.farZ = c.cam[13];
        }
        d.W = v.width; d.H = v.height; d.spp = v.spp; d.frame = v.frame;
        d.scene = si;
        d.pixBase = uint32_t(nPix);
        d.pathBase = uint32_t(nPaths);
        d.out = reinterpret_cast<float4*>(v.d_color);
        nPix += uint64_t(v.width) * v.height;
        nPaths += uint64_t(v.width) * v.height * v.spp;
        if (nPaths > 0xffffffffull / 3) return fail(FHP_EINVAL, "batch too large");
    }
    std::vector<SceneDev> stab(scenes.size());
    for (size_t i = 0; i < scenes.size(); i++) stab[i] = scenes[i]->dev;
    int rc = ensurePassBuffersN(p, nPaths, nPix);
    if (rc) return rc;
    HIP_OR_FAIL(p->viewTab.alloc(vt.size() * sizeof(ViewDev)));
    HIP_OR_FAIL(p->sceneTab.alloc(stab.size() * sizeof(SceneDev)));
    // Tables go through pinned staging memory, so the copies are asynchronous and the call does
    // not wait for the batch; the previous call's copies must have left the staging buffer first.
    const size_t vb = vt.size() * sizeof(ViewDev), sb = stab.size() * sizeof(SceneDev);
    if (p->stageDone) HIP_OR_FAIL(hipEventSynchronize(p->stageDone));
    else HIP_OR_FAIL(hipEventCreateWithFlags(&p->stageDone, hipEventDisableTiming));
    if (p->stageBytes < vb + sb) {
        if (p->stage) HIP_OR_FAIL(hipHostFree(p->stage));
        p->stage = nullptr;
        p->stageBytes = 0;
        HIP_OR_FAIL(hipHostMalloc(&p->stage, vb + sb, hipHostMallocDefault));
        p->stageBytes = vb + sb;
    }
    std::memcpy(p->stage, vt.data(), vb);
    std::memcpy(static_cast<uint8_t*>(p->stage) + vb, stab.data(), sb);
    HIP_OR_FAIL(hipMemcpyAsync(p->viewTab.p, p->stage, vb, hipMemcpyHostToDevice, st));
    HIP_OR_FAIL(hipMemcpyAsync(p->sceneTab.p, static_cast<uint8_t*>(p->stage) + vb, sb, hipMemcpyHostToDevice, st));
    HIP_OR_FAIL(hipEventRecord(p->stageDone, st));
    const bool count = p->get("countTraversal") != 0;
    unsigned long long* dstats = p->stats.as<unsigned long long>();
    if (h_stats) HIP_OR_FAIL(launchFill(reinterpret_cast<float*>(dstats), 16, 0.f, st));
    MptArgs a;
    std::memset(&a, 0, sizeof(a));  // padding too: the bytes key the captured graph
    a.maxBounces = uint32_t(p->get("maxBounces"));
    a.computeDirect = p->get("computeDirect") != 0;
    a.importance = p->get("useImportanceSampling") != 0;
    a.nee = p->get("useEmissiveNEE") != 0;
    a.lightKind = uint32_t(p->get("emissiveSampler"));
    a.envNee = p->get("useEnvNEE") != 0;
    if (a.nee && !a.computeDirect) return fail(FHP_EINVAL, "useEmissiveNEE requires computeDirect");
    if (a.envNee && !a.computeDirect) return fail(FHP_EINVAL, "useEnvNEE requires computeDirect");
    bool otherLights = false;
    for (FhpScene sc : scenes) otherLights |= sc->dev.nALights || (a.envNee && sc->dev.envImp && (sc->dev.envFlags & kEnvLight));
    a.shadowRays = (a.nee || otherLights) ? 1u : 0u;
    a.lightMode = otherLights ? 2u : (a.nee ? 1u : 0u);
    a.separateShadow = 0;
    a.tailPaths = uint32_t(p->get("tailPaths"));
    a.tailLanes = uint32_t(p->get("tailLanes"));
    a.nPaths = uint32_t(nPaths);
    a.nPix = uint32_t(nPix);
    for (FhpScene sc : scenes) a.keepMisses |= (sc->dev.envTex && (sc->dev.envFlags & kEnvLight)) ? 1u : 0u;
    a.views = p->viewTab.as<ViewDev>();
    a.scenes = p->sceneTab.as<SceneDev>();
    a.nViews = nv;
    HIP_OR_FAIL(p->pathView.alloc(size_t(nPaths) * 2));
    a.pathView = p->pathView.as<uint16_t>();
    a.thp = p->thp.as<float4>(); a.prevN = p->prevN.as<float4>(); a.rng = p->rng.as<uint4>();
    a.Lp = p->Lp.as<float4>(); a.Ls = p->Ls.as<float4>();
    a.rayQ[0] = p->rayQ0.as<float4>(); a.rayQ[1] = p->rayQ1.as<float4>();
    a.pathQ[0] = p->pathQ0.as<uint32_t>(); a.pathQ[1] = p->pathQ1.as<uint32_t>();
    a.hitQ = p->hitQ.as<uint4>(); a.hitIdx = p->hitIdx.as<uint32_t>(); a.shQ = p->shQ.as<float4>();
    a.pixHit = p->pixHit.as<uint4>();
    a.counters = p->counters.as<uint32_t>();
    if ((rc = prepareTimer(p))) return rc;
    TraceScratch ts;
    if ((rc = passTrace(deepest, p, ts))) return rc;
    const bool graph = p->get("useGraph") != 0 && !count && !p->timer.on;
    if (!graph) {
        HIP_OR_FAIL(launchMptBatch(a, ts, count, h_stats ? dstats : nullptr, st, &p->timer));
    } else {
        // Captured bounce loop, keyed by the arguments and the traversal configuration (the
        // tables' contents are read from device memory at run time).
        std::vector<uint8_t> key(sizeof(MptArgs) + sizeof(TraceScratch));
        std::memcpy(key.data(), &a, sizeof(MptArgs));
        std::memcpy(key.data() + sizeof(MptArgs), &ts, sizeof(TraceScratch));
        if (key != p->batchKey) {
            if (p->batchGraph) (void)hipGraphExecDestroy(p->batchGraph);
            p->batchGraph = nullptr;
            p->batchKey.clear();
            if (!p->capStream) HIP_OR_FAIL(hipStreamCreateWithFlags(&p->capStream, hipStreamNonBlocking));
            hipGraph_t g = nullptr;
            HIP_OR_FAIL(hipStreamBeginCapture(p->capStream, hipStreamCaptureModeThreadLocal));
            hipError_t ce = launchMptBatchBounces(a, ts, false, nullptr, p->capStream, nullptr);
            hipError_t ee = hipStreamEndCapture(p->capStream, &g);
            HIP_OR_FAIL(ce);
            HIP_OR_FAIL(ee);
            hipError_t ie = hipGraphInstantiate(&p->batchGraph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HIP_OR_FAIL(ie);
            p->batchKey = key;
        }
        HIP_OR_FAIL(hipGraphLaunch(p->batchGraph, st));
        HIP_OR_FAIL(launchMptBatchEpilogue(a, h_stats ? dstats : nullptr, st));
    }
    if ((rc = collectTimer(p, st))) return rc;
    if (h_stats) {
        unsigned long long hs[8];
        HIP_OR_FAIL(hipMemcpyAsync(hs, dstats, 64, hipMemcpyDeviceToHost, st));
        HIP_OR_FAIL(hipStreamSynchronize(st));
        for (int i = 0; i < 8; i++) h_stats[i] = hs[i];
    }
    return FHP_OK;
}

// ------------------------------------------------------------------ gradients
int fhp_grads_create(FhpScene s, uint32_t hashSize, int withTex, FhpGrads* out) {
    if (!s || !out) return fail(FHP_EINVAL, "null argument");
    if (hashSize == 0 || hashSize > 65536) return fail(FHP_EINVAL, "hash size must be in [1, 65536]");
    std::unique_ptr<FhpGrads_> g(new FhpGrads_());
    g->scene = s;
    g->hashSize = hashSize;
    g->dim = s->h.nM * 20;
    HIP_OR_FAIL(g->tmp.alloc(size_t(g->dim) * hashSize * 4));
    HIP_OR_FAIL(g->agg.alloc(size_t(g->dim) * 4));
    g->withTex = withTex != 0 && !s->h.texels.empty();
    if (g->withTex) HIP_OR_FAIL(g->tex.alloc(s->h.texels.size() * 4));
    *out = g.release();
    return fhp_grads_clear(*out, nullptr);
}
int fhp_grads_destroy(FhpGrads g) {
    delete g;
    return FHP_OK;
}
int fhp_grads_clear(FhpGrads g, FhpStream stream) {
    if (!g) return fail(FHP_EINVAL, "null argument");
    hipStream_t st = hipStream_t(stream);
    HIP_OR_FAIL(launchFill(g->tmp.as<float>(), uint64_t(g->dim) * g->hashSize, 0.f, st));
    HIP_OR_FAIL(launchFill(g->agg.as<float>(), g->dim, 0.f, st));
    if (g->withTex) HIP_OR_FAIL(launchFill(g->tex.as<float>(), g->scene->h.texels.size(), 0.f, st));
    if (g->withMesh) {
        const uint64_t n = 3ull * g->meshVerts;
        HIP_OR_FAIL(launchFill(g->posTmp.as<float>(), n * g->hashSize, 0.f, st));
        HIP_OR_FAIL(launchFill(g->nrmTmp.as<float>(), n * g->hashSize, 0.f, st));
        HIP_OR_FAIL(launchFill(g->posAgg.as<float>(), n, 0.f, st));
        HIP_OR_FAIL(launchFill(g->nrmAgg.as<float>(), n, 0.f, st));
    }
    return FHP_OK;
}
int fhp_grads_aggregate(FhpGrads g, FhpStream stream) {
    if (!g) return fail(FHP_EINVAL, "null argument");
    HIP_OR_FAIL(launchAggregate(g->tmp.as<float>(), g->hashSize, g->dim, g->agg.as<float>(), hipStream_t(stream)));
    if (g->withMesh) {
        HIP_OR_FAIL(launchAggregate(g->posTmp.as<float>(), g->hashSize, 3 * g->meshVerts, g->posAgg.as<float>(),
                                    hipStream_t(stream)));
        HIP_OR_FAIL(launchAggregate(g->nrmTmp.as<float>(), g->hashSize, 3 * g->meshVerts, g->nrmAgg.as<float>(),
                                    hipStream_t(stream)));
    }
    return FHP_OK;
}
int fhp_grads_set_mesh(FhpGrads g, uint32_t mesh) {
    if (!g) return fail(FHP_EINVAL, "null argument");
    const HostScene& h = g->scene->h;
    if (mesh + 1 >= h.meshOff.size()) return fail(FHP_EINVAL, "mesh id out of range");
    g->mesh = mesh;
    g->meshV0 = h.meshOff[mesh];
    g->meshVerts = h.meshOff[mesh + 1] - h.meshOff[mesh];
    const size_t n = 3ull * g->meshVerts * 4;
    HIP_OR_FAIL(g->posTmp.alloc(n * g->hashSize));
    HIP_OR_FAIL(g->nrmTmp.alloc(n * g->hashSize));
    HIP_OR_FAIL(g->posAgg.alloc(n));
    HIP_OR_FAIL(g->nrmAgg.alloc(n));
    g->withMesh = true;
    return fhp_grads_clear(g, nullptr);
}
int fhp_grads_get(FhpGrads g, float** dm, uint64_t* nm, float** dt, uint64_t* nt) {
    if (!g) return fail(FHP_EINVAL, "null argument");
    if (dm) *dm = g->agg.as<float>();
    if (nm) *nm = g->dim;
    if (dt) *dt = g->withTex ? g->tex.as<float>() : nullptr;
    if (nt) *nt = g->withTex ? g->scene->h.texels.size() : 0;
    return FHP_OK;
}

int fhp_grads_copy(FhpGrads g, int which, float* dst, uint64_t n, FhpStream stream) {
    if (!g || !dst) return fail(FHP_EINVAL, "null argument");
    if (which < 0 || which > 3) return fail(FHP_EINVAL, "which must be 0 (material), 1 (texture), 2 (mesh position) or 3 (mesh normal)");
    const void* src = which == 0 ? g->agg.p : which == 1 ? g->tex.p : which == 2 ? g->posAgg.p : g->nrmAgg.p;
    const uint64_t have = which == 0   ? g->dim
                          : which == 1 ? (g->withTex ? g->scene->h.texels.size() : 0)
                                       : (g->withMesh ? 3ull * g->meshVerts : 0);
    if (n != have || have == 0) return fail(FHP_EINVAL, "gradient size mismatch");
    HIP_OR_FAIL(hipMemcpyAsync(dst, src, n * 4, hipMemcpyDeviceToDevice, hipStream_t(stream)));
    return FHP_OK;
}

namespace {
// A batched view resolved for the kernels (camera from 14 floats, or the scene's).
WarView warViewFrom(const FhpScene_* s, const float* h_camera, uint32_t frame, const float* dLdI, float* out) {
    WarView v{};
    v.cam = s->dev.cam;
    if (h_camera) {
        HostScene c;
        std::memcpy(c.cam, h_camera, sizeof(c.cam));
        computeCamera(c);
        v.cam.pos = mk3(c.cam[0], c.cam[1], c.cam[2]);
        v.cam.U = mk3(c.camU[0], c.camU[1], c.camU[2]);
        v.cam.V = mk3(c.camV[0], c.camV[1], c.camV[2]);
        v.cam.W = mk3(c.camW[0], c.camW[1], c.camW[2]);
        v.cam.nearZ = c.cam[12];
        v.cam.farZ = c.cam[13];
    }
    v.frame = frame;
    v.dLdI = dLdI;
    v.out = reinterpret_cast<float4*>(out);
    return v;
}

}  // namespace

// WARDiffPathTracer views of one scene and one frame shape rendered in one launch sequence: the
// primal of every view (Primal, BackwardDiff and ForwardDiffDebug without a debug variable),
// resolved into each view's d_color; BackwardDiffDebug clears them.
static int warExecuteBatch(FhpPass p, uint32_t nv, const FhpView* views, uint64_t* h_stats, FhpStream stream) {
    FhpScene s = views[0].scene;
    const uint32_t W = views[0].width, H = views[0].height, spp = views[0].spp;
    for (uint32_t i = 0; i < nv; i++) {
        const FhpView& v = views[i];
        if (!v.scene || !v.d_color) return fail(FHP_EINVAL, "view " + std::to_string(i) + ": null scene or output");
        if (v.scene != s || v.width != W || v.height != H || v.spp != spp)
            return fail(FHP_EINVAL, "batched WARDiffPathTracer views share one scene, width, height and spp");
    }
    if (W == 0 || H == 0 || spp == 0 || W > 65536 || H > 65536) return fail(FHP_EINVAL, "bad frame shape");
    const uint64_t n = uint64_t(W) * H * spp * nv;
    if (n > 0xffffffffull / 3) return fail(FHP_EINVAL, "batch too large");
    const int mode = int(p->get("diffMode"));
    if (mode == kDiffForwardDebug && p->dbgVar != kDiffVarNone)
        return fail(FHP_EINVAL, "batched ForwardDiffDebug gradient images are not supported");
    hipStream_t st = hipStream_t(stream);
    const uint64_t npix = uint64_t(W) * H;
    if (mode == kDiffBackwardDebug) {
        for (uint32_t i = 0; i < nv; i++) HIP_OR_FAIL(launchFill(views[i].d_color, npix * 4, 0.f, st));
        return FHP_OK;
    }
    int rc = ensurePassBuffersN(p, n, npix);
    if (rc) return rc;
    unsigned long long* dstats = p->stats.as<unsigned long long>();
    if (h_stats) HIP_OR_FAIL(launchFill(reinterpret_cast<float*>(dstats), 16, 0.f, st));
    const bool fixed = p->get("useFixedSeed") != 0;
    std::vector<WarView> vt(nv);
    for (uint32_t i = 0; i < nv; i++)
        vt[i] = warViewFrom(s, views[i].h_camera, fixed ? uint32_t(p->get("fixedSeed")) : views[i].frame * spp, nullptr,
                            views[i].d_color);
    HIP_OR_FAIL(p->warViewTab.alloc(vt.size() * sizeof(WarView)));
    HIP_OR_FAIL(hipMemcpyAsync(p->warViewTab.p, vt.data(), vt.size() * sizeof(WarView), hipMemcpyHostToDevice, st));
    HIP_OR_FAIL(hipStreamSynchronize(st));
    TraceScratch ts;
    if ((rc = passTrace(s, p, ts))) return rc;
    WarArgs a{};
    a.W = W; a.H = H; a.spp = spp;
    a.maxBounces = uint32_t(p->get("maxBounces"));
    a.useBSDFSampling = p->get("useBSDFSampling") != 0;
    a.useNEE = p->get("useNEE") != 0;
    a.useMIS = p->get("useMIS") != 0;
    a.useAntithetic = p->get("useAntitheticSampling") != 0;
    a.tmp = p->Lp.as<float4>();
    a.rays = h_stats ? dstats : nullptr;
    a.views = p->warViewTab.as<WarView>();
    a.nViews = nv;
    if ((rc = prepareTimer(p))) return rc;
    HIP_OR_FAIL(launchWarPrimal(s->dev, a, ts, st, &p->timer));
    if ((rc = collectTimer(p, st))) return rc;
    if (h_stats) {
        unsigned long long hs[8];
        HIP_OR_FAIL(hipMemcpyAsync(hs, dstats, 64, hipMemcpyDeviceToHost, st));
        HIP_OR_FAIL(hipStreamSynchronize(st));
        for (int i = 0; i < 8; i++) h_stats[i] = hs[i];
    }
    return FHP_OK;
}

namespace {
// The WARDiffPathTracer adjoint (fhp_pass_backward) of one frame or, with views, of a batch of
// views of the same shape in one launch sequence. The view table is staged from host memory; the
// stream is synchronised before returning when views are given.
int warBackward(FhpPass p, FhpScene s, uint32_t W, uint32_t H, uint32_t spp, uint32_t frame, const float* d_dLdI,
                const std::vector<WarView>* views, FhpGrads g, uint64_t* h_stats, FhpStream stream) {
    if (!p || !s || !g) return fail(FHP_EINVAL, "null argument");
    if (p->kind != "WARDiffPathTracer") return fail(FHP_ESTATE, "backward needs a WARDiffPathTracer pass");
    // runBackward runs the adjoint in BackwardDiff mode (WARDiffPathTracer.rt.slang:161-180) and, with
    // dL/dI = (1, 0, 0) at every pixel, in BackwardDiffDebug mode (:201-211).
    const int mode = int(p->get("diffMode"));
    if (mode != kDiffBackward && mode != kDiffBackwardDebug)
        return fail(FHP_ESTATE, "backward needs diffMode BackwardDiff (1) or BackwardDiffDebug (3)");
    const uint32_t nviews = views ? uint32_t(views->size()) : 1u;
    if (mode == kDiffBackward && !views && !d_dLdI) return fail(FHP_EINVAL, "null argument");
    if (mode == kDiffBackward && views)
        for (uint32_t i = 0; i < nviews; i++)
            if (!(*views)[i].dLdI) return fail(FHP_EINVAL, "view " + std::to_string(i) + ": null dL/dI");
    if (g->scene != s) return fail(FHP_EINVAL, "gradients belong to another scene");
    if (W == 0 || H == 0 || spp == 0) return fail(FHP_EINVAL, "empty frame");
    if (W > 65536 || H > 65536) return fail(FHP_EINVAL, "frame dimension above 65536");
    if (uint64_t(W) * H * spp * nviews > 0xffffffffull / 3)
        return fail(FHP_EINVAL, "frame too large (width * height * spp paths)");
    const bool war = p->get("useWAR") != 0;
    if (war) {
        if (p->get("maxBounces") > kMaxGeomBounces)
            return fail(FHP_EINVAL, "useWAR=1 (vertex gradients) supports maxBounces up to " + std::to_string(kMaxGeomBounces));
        if (s->dev.hasTransmission)
            return fail(FHP_EINVAL, "useWAR=1 (vertex gradients) supports reflective materials only (no transmission lobes)");
        if (!g->withMesh) return fail(FHP_ESTATE, "useWAR=1 needs mesh gradients (fhp_grads_set_mesh)");
        const double aux = p->get("auxSampleCount");
        if (aux < 1 || aux > 64 || double(uint32_t(aux)) != aux) return fail(FHP_EINVAL, "auxSampleCount must be in [1, 64]");
        if (p->get("useAntitheticSampling") != 0 && (uint32_t(aux) & 1u))
            return fail(FHP_EINVAL, "auxSampleCount must be even with useAntitheticSampling");
    }
    hipStream_t st = hipStream_t(stream);
    int rc = ensurePassBuffers(p, 1, 1, 1);
    if (rc) return rc;
    unsigned long long* dstats = p->stats.as<unsigned long long>();
    if (h_stats) HIP_OR_FAIL(launchFill(reinterpret_cast<float*>(dstats), 12, 0.f, st));
    WarArgs a{};
    a.W = W; a.H = H; a.spp = spp;
    a.frame = p->get("useFixedSeed") != 0 ? uint32_t(p->get("fixedSeed")) : frame * spp;
    a.maxBounces = uint32_t(p->get("maxBounces"));
    a.useBSDFSampling = p->get("useBSDFSampling") != 0;
    a.useNEE = p->get("useNEE") != 0;
    a.useMIS = p->get("useMIS") != 0;
    a.useAntithetic = p->get("useAntitheticSampling") != 0;
    a.dLdI = mode == kDiffBackwardDebug ? nullptr : d_dLdI;
    if (views) {
        std::vector<WarView> vt = *views;
        if (mode == kDiffBackwardDebug)
            for (WarView& v : vt) v.dLdI = nullptr;
        HIP_OR_FAIL(p->warViewTab.alloc(vt.size() * sizeof(WarView)));
        HIP_OR_FAIL(hipMemcpyAsync(p->warViewTab.p, vt.data(), vt.size() * sizeof(WarView), hipMemcpyHostToDevice, st));
        // The staging copy reads vt: wait for it before vt goes out of scope.
        HIP_OR_FAIL(hipStreamSynchronize(st));
        a.views = p->warViewTab.as<WarView>();
        a.nViews = nviews;
    }
    a.gradMat = g->tmp.as<float>();
    a.hashSize = g->hashSize;
    a.matGradDim = g->dim;
    a.gradTex = g->withTex ? g->tex.as<float>() : nullptr;
    a.rays = h_stats ? dstats : nullptr;
    if ((rc = prepareTimer(p))) return rc;
    TraceScratch ts;
    if ((rc = passTrace(s, p, ts))) return rc;
    if ((rc = prepareHitLog(p, s, ts, a))) return rc;
    HIP_OR_FAIL(launchWarAdjoint(s->dev, a, ts, st, &p->timer));
    if (war) {
        GeomArgs ga{};
        ga.W = W; ga.H = H; ga.spp = spp; ga.frame = a.frame;
        ga.useBSDFSampling = a.useBSDFSampling; ga.useNEE = a.useNEE; ga.useMIS = a.useMIS;
        ga.useAntithetic = a.useAntithetic;
        ga.aux = uint32_t(p->get("auxSampleCount"));
        ga.kappa = float(std::pow(10.0, p->get("Log10vMFConcentration")));
        ga.kappaScreen = float(std::pow(10.0, p->get("Log10vMFConcentrationScreen")));
        ga.neeHeld = p->get("useSolidAngleNEEWarp") != 0;
        ga.dLdI = a.dLdI;
        ga.gradPos = g->posTmp.as<float>();
        ga.gradNrm = g->nrmTmp.as<float>();
        ga.hashSize = g->hashSize; ga.meshV0 = g->meshV0; ga.meshVerts = g->meshVerts;
        ga.rays = h_stats ? dstats + 2 : nullptr;
        ga.views = a.views;
        ga.nViews = a.nViews;
        // The kernel's grid has at most ts.gridBlocks blocks of ts.block threads.
        if (p->get("useHitLog") != 0) {
            HIP_OR_FAIL(p->auxLog.alloc(size_t(3) * ga.aux * ts.gridBlocks * ts.block * 4));
            ga.auxLog = p->auxLog.as<uint32_t>();
        }
        ga.maxBounces = a.maxBounces;
        if (a.maxBounces > 0) {
            // maxBounces = 0 runs the one-pass direct-illumination instances, which keep no records.
            HIP_OR_FAIL(p->vtxLog.alloc(size_t(kGeomVertexWords) * (a.maxBounces + 2) * ts.gridBlocks * ts.block * 4));
            ga.vtxLog = p->vtxLog.as<uint32_t>();
        }
        HIP_OR_FAIL(launchWarGeom(s->dev, ga, ts, st, &p->timer));
    }
    if ((rc = collectTimer(p, st))) return rc;
    if (h_stats) {
        unsigned long long hs[6];
        HIP_OR_FAIL(hipMemcpyAsync(hs, dstats, 48, hipMemcpyDeviceToHost, st));
        HIP_OR_FAIL(hipStreamSynchronize(st));
        for (int i = 0; i < 6; i++) h_stats[i] = hs[i];
    }
    return FHP_OK;
}
}  // namespace

int fhp_pass_backward(FhpPass p, FhpScene s, uint32_t W, uint32_t H, uint32_t spp, uint32_t frame, const float* d_dLdI,
                      FhpGrads g, uint64_t* h_stats, FhpStream stream) {
    return warBackward(p, s, W, H, spp, frame, d_dLdI, nullptr, g, h_stats, stream);
}

int fhp_pass_backward_batch(FhpPass p, FhpScene s, uint32_t nv, const FhpDiffView* views, uint32_t W, uint32_t H,
                            uint32_t spp, FhpGrads g, uint64_t* h_stats, FhpStream stream) {
    if (!p || !s || (nv && !views)) return fail(FHP_EINVAL, "null argument");
    if (nv == 0) return FHP_OK;
    std::vector<WarView> vt(nv);
    const bool fixed = p->get("useFixedSeed") != 0;
    for (uint32_t i = 0; i < nv; i++)
        vt[i] = warViewFrom(s, views[i].h_camera, fixed ? uint32_t(p->get("fixedSeed")) : views[i].frame * spp,
                            views[i].d_dLdI, nullptr);
    return warBackward(p, s, W, H, spp, 0, nullptr, &vt, g, h_stats, stream);
}

// ------------------------------------------------------------------ AccumulatePass
int fhp_accumulate(FhpPass p, const void* d_in, int fmt, uint32_t W, uint32_t H, float* d_out, uint32_t refresh,
                   FhpStream stream) {
    if (!p || !d_in || !d_out) return fail(FHP_EINVAL, "null argument");
    if (p->kind != "AccumulatePass") return fail(FHP_EINVAL, "fhp_accumulate needs an AccumulatePass");
    if (fmt < kAccumInputFloat || fmt > kAccumInputSint) return fail(FHP_EINVAL, "input format: 0 float, 1 uint, 2 int");
    if (W == 0 || H == 0) return fail(FHP_EINVAL, "empty frame");
    if (W > 65536 || H > 65536) return fail(FHP_EINVAL, "frame dimension above 65536");
    hipStream_t st = hipStream_t(stream);
    // AccumulatePass::execute (AccumulatePass.cpp:127-218).
    if (p->get("autoReset") != 0 && refresh) p->accCount = 0;
    const uint32_t maxFrames = uint32_t(p->get("maxFrameCount"));
    const int overflow = int(p->get("overflowMode"));
    if (maxFrames > 0 && p->accCount == maxFrames) {
        if (overflow == 0) return FHP_OK;  // Stop: the output keeps the accumulated image
        if (overflow == 1) p->accCount = 0;  // Reset; EMA continues with weight 1 / (maxFrameCount + 1)
    }
    if (W != p->accW || H != p->accH) {
        p->accW = W;
        p->accH = H;
        p->accCount = 0;
    }
    const uint64_t n = uint64_t(W) * H;
    const bool enabled = p->get("enabled") != 0;
    if (!enabled && fmt == kAccumInputFloat) {
        // Disabled with float input: the input is copied through (the reference blits it).
        HIP_OR_FAIL(hipMemcpyAsync(d_out, d_in, n * 16, hipMemcpyDeviceToDevice, st));
        return FHP_OK;
    }
    // AccumulatePass::accumulate (:220-292) and prepareAccumulation (:380-420): the sums of the
    // current precision mode are (re)created on a mode or size change, which restarts the
    // accumulation; a frame count of 0 treats them as cleared.
    const int mode = int(p->get("precisionMode"));
    const uint64_t sumBytes = n * (mode == kAccumDouble ? 32 : 16);
    if (mode != p->accMode || p->accSum.bytes != sumBytes) {
        p->accSum.release();
        p->accCorr.release();
        HIP_OR_FAIL(p->accSum.alloc(sumBytes));
        if (mode == kAccumSingleCompensated) HIP_OR_FAIL(p->accCorr.alloc(n * 16));
        p->accMode = mode;
        p->accCount = 0;
    }
    AccumArgs a{};
    a.in = d_in;
    a.out = reinterpret_cast<float4*>(d_out);
    a.sum = mode == kAccumDouble ? nullptr : p->accSum.as<float4>();
    a.corr = mode == kAccumSingleCompensated ? p->accCorr.as<float4>() : nullptr;
    a.sumD = mode == kAccumDouble ? p->accSum.as<double2>() : nullptr;
    a.n = n;
    a.count = p->accCount;
    a.flags = (enabled ? kAccumEnabled : 0u) | (maxFrames > 0 ? kAccumMovingAverage : 0u) | (p->accCount == 0 ? kAccumFirst : 0u);
    a.mode = mode;
    a.inputFormat = fmt;
    HIP_OR_FAIL(launchAccumulate(a, st));
    // The frame limit does not apply to SingleCompensated (:279-284).
    if (maxFrames == 0 || mode == kAccumSingleCompensated || p->accCount < maxFrames) p->accCount++;
    return FHP_OK;
}
int fhp_accumulate_reset(FhpPass p) {
    if (!p) return fail(FHP_EINVAL, "null argument");
    if (p->kind != "AccumulatePass") return fail(FHP_EINVAL, "not an AccumulatePass");
    p->accCount = 0;
    return FHP_OK;
}
int fhp_accumulate_frame_count(FhpPass p, uint32_t* count) {
    if (!p || !count) return fail(FHP_EINVAL, "null argument");
    if (p->kind != "AccumulatePass") return fail(FHP_EINVAL, "not an AccumulatePass");
    *count = p->accCount;
    return FHP_OK;
}

// ------------------------------------------------------------------ ImageCompare
int fhp_image_compare(int metric, const float* d_a, const float* d_b, uint32_t W, uint32_t H, int alpha, float* d_error_map,
                      float* d_heat_map, double* h_error, FhpStream stream) {
    if (!d_a || !d_b || !h_error) return fail(FHP_EINVAL, "null argument");
    if (metric < kMetricMSE || metric > kMetricMAPE) return fail(FHP_EINVAL, "metric: 0 mse, 1 rmse, 2 mae, 3 mape");
    if (W == 0 || H == 0) return fail(FHP_EINVAL, "empty image");
    if (d_heat_map && !d_error_map) return fail(FHP_EINVAL, "a heat map needs the error map");
    hipStream_t st = hipStream_t(stream);
    const uint64_t n = uint64_t(W) * H;
    DevBuf scratch;
    HIP_OR_FAIL(scratch.alloc(size_t(comparePartials(n) + 1) * sizeof(CompareResult)));
    CompareResult* part = scratch.as<CompareResult>();
    CompareResult* res = part + comparePartials(n);
    HIP_OR_FAIL(launchCompare(metric, reinterpret_cast<const float4*>(d_a), reinterpret_cast<const float4*>(d_b), n,
                              alpha ? 4 : 3, d_error_map, part, res, st));
    if (d_heat_map) HIP_OR_FAIL(launchHeatMap(d_error_map, n, res, reinterpret_cast<float4*>(d_heat_map), st));
    CompareResult r{};
    HIP_OR_FAIL(hipMemcpyAsync(&r, res, sizeof(r), hipMemcpyDeviceToHost, st));
    HIP_OR_FAIL(hipStreamSynchronize(st));
    *h_error = r.sum / double(n);
    return FHP_OK;
}

// ------------------------------------------------------------------ V-buffer
int fhp_pass_set_output(FhpPass p, const char* channel, void* d_data) {
    if (!p || !channel) return fail(FHP_EINVAL, "null argument");
    if (p->kind != "WARDiffPathTracer") return fail(FHP_EINVAL, "output channels are bound on a WARDiffPathTracer");
    const std::string c(channel);
    if (c == "dColor") p->outDColor = static_cast<float4*>(d_data);
    else return fail(FHP_EINVAL, "unknown output channel '" + c + "' (dColor)");
    return FHP_OK;
}

int fhp_pass_set_diff_debug(FhpPass p, int var_type, uint32_t id, uint32_t offset, const float* h_grad) {
    if (!p || (var_type != kDiffVarNone && !h_grad)) return fail(FHP_EINVAL, "null argument");
    if (p->kind != "WARDiffPathTracer") return fail(FHP_EINVAL, "debug parameters belong to a WARDiffPathTracer");
    if (var_type < kDiffVarNone || var_type > kDiffVarGeometryTranslation)
        return fail(FHP_EINVAL, "var_type: 0 None, 1 Material, 2 GeometryTranslation");
    if (var_type == kDiffVarMaterial && offset >= 20) return fail(FHP_EINVAL, "material parameter offset must be below 20");
    p->dbgVar = var_type;
    p->dbgId = id;
    p->dbgOffset = offset;
    for (int i = 0; i < 4; i++) p->dbgGrad[i] = var_type == kDiffVarNone ? 0.f : h_grad[i];
    return FHP_OK;
}

int fhp_pass_set_input(FhpPass p, const char* channel, const void* d_data) {
    if (!p || !channel) return fail(FHP_EINVAL, "null argument");
    if (p->kind != "MinimalPathTracer") return fail(FHP_EINVAL, "input channels are bound on a MinimalPathTracer");
    const std::string c(channel);
    if (c == "vbuffer") p->inVbuf = static_cast<const uint32_t*>(d_data);
    else if (c == "viewW") p->inViewW = static_cast<const float4*>(d_data);
    else return fail(FHP_EINVAL, "unknown input channel '" + c + "' (vbuffer, viewW)");
    return FHP_OK;
}

int fhp_vbuffer_execute(FhpPass p, FhpScene s, uint32_t W, uint32_t H, uint32_t* d_vbuffer, float* d_depth, float* d_viewW,
                        float* d_mask, float* h_jitter, FhpStream stream) {
    if (!p || !s || !d_vbuffer) return fail(FHP_EINVAL, "null argument");
    if (p->kind != "VBufferRT") return fail(FHP_EINVAL, "fhp_vbuffer_execute needs a VBufferRT pass");
    if (W == 0 || H == 0) return fail(FHP_EINVAL, "empty frame");
    if (W > 65536 || H > 65536 || uint64_t(W) * H > 0xffffffffull) return fail(FHP_EINVAL, "frame too large");
    if (!s->h.meshTriContiguous) return fail(FHP_ESTATE, "V-buffer encoding needs each mesh's triangles to be contiguous");
    // Camera::beginFrame (Camera.cpp:51-58): one jitter sample per frame, scaled by 1 / frame size.
    float jx, jy;
    p->jitter.next(jx, jy);
    jx *= 1.f / float(W);
    jy *= 1.f / float(H);
    if (h_jitter) {
        h_jitter[0] = jx;
        h_jitter[1] = jy;
    }
    TraceScratch ts;
    if (int rc = passTrace(s, p, ts)) return rc;
    HIP_OR_FAIL(launchVBuffer(s->dev, W, H, jx, jy, d_vbuffer, d_depth, reinterpret_cast<float4*>(d_viewW), d_mask, ts,
                              hipStream_t(stream)));
    return FHP_OK;
}

}  // extern "C"
