// WARDiffPathTracer geometry-gradient kernel (warped-area reparameterization); the per-path code
// is in fhp_war_geom.h. DIRECT instances run maxBounces = 0 paths in one pass, the others record
// each path's vertices and sweep them in reverse.
#include "fhp_war_geom.h"

namespace fhp {

namespace {

// Unconstrained the kernel takes 248 VGPRs (2 waves per SIMD); capped at 128 (4 waves) it spills
// more but runs the 256x256, 8 spp, 16-aux vertex-gradient pass about 1.2x faster.
constexpr int kGeomWavesPerEU = 4;

template <int B, int D, bool DIRECT>
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(kGeomWavesPerEU))) void kWarGeom(SceneDev s, GeomArgs a,
                                                                                                     TraceScratch ts) {
    __shared__ uint32_t lds[D * B];
    TraceStack st = makeStack<B, D>(lds, ts, blockIdx.x, threadIdx.x);
    const uint32_t n = a.W * a.H * a.spp * (a.views ? a.nViews : 1u);
    const uint32_t G = gridDim.x * B;
    unsigned long long nc = 0;
    // Aux-hit log of this thread: 3 * aux words [word][thread] in a.auxLog (3 * aux * G words).
    const AuxLog lg = a.auxLog ? AuxLog{(GlobalU32*)(a.auxLog + blockIdx.x * B + threadIdx.x), G} : noAuxLog();
    // Vertex records of this thread: kGeomVertexWords * (maxBounces + 2) words [word][thread].
    const VertexLog vl{(GlobalU32*)(a.vtxLog + blockIdx.x * B + threadIdx.x), G, a.maxBounces + 2};
    for (uint32_t id = blockIdx.x * B + threadIdx.x; id < n; id += G) geomItem<DIRECT>(s, a, id, st, nc, lg, vl);
    if (a.rays) atomicAdd(&a.rays[0], nc);
}

}  // namespace

hipError_t launchWarGeom(const SceneDev& s, const GeomArgs& a, const TraceScratch& ts, hipStream_t st, KernelTimer* tm) {
    const uint32_t n = a.W * a.H * a.spp * (a.views ? a.nViews : 1u);
    if (tm) tm->begin(KK_WAR_ADJOINT, st);
    hipError_t e = withTraceConfig(ts, [&](auto Bc, auto Dc) {
        constexpr int B = decltype(Bc)::value, D = decltype(Dc)::value;
        // Persistent grid: no more blocks than are resident at kGeomWavesPerEU waves per SIMD.
        const uint32_t resident = uint32_t(numCUs()) * std::max<uint32_t>(1u, kGeomWavesPerEU * 4 * 64 / B);
        const uint32_t blocks = std::min<uint32_t>(std::min<uint32_t>(ts.gridBlocks, resident), (n + B - 1) / B);
        if (a.maxBounces == 0) hipLaunchKernelGGL((kWarGeom<B, D, true>), dim3(blocks), dim3(B), 0, st, s, a, ts);
        else hipLaunchKernelGGL((kWarGeom<B, D, false>), dim3(blocks), dim3(B), 0, st, s, a, ts);
        return hipGetLastError();
    });
    if (tm) tm->end(st);
    return e;
}

}  // namespace fhp
