// WARDiffPathTracer megakernels (one path per thread), shared by the translation units that
// instantiate them (fhp_war.hip: instances without the transmission adjoint; fhp_war_trans.hip:
// the instances with it, for scenes with transmissive materials).
#pragma once
#include "fhp_war_path.h"

namespace fhp {
namespace {

// The megakernels hold a whole path (and its adjoint) per lane. Capped at 96 VGPRs (5 waves per
// SIMD) they spill some state to scratch; unconstrained (3 / 2 waves per SIMD) latency hiding
// suffers more (the cap measured 1.47x faster per 512x512, 16 spp forward+adjoint iteration).
#ifndef FHP_WAR_WAVES_PER_EU
#define FHP_WAR_WAVES_PER_EU 5
#endif
constexpr int kWarWavesPerEU = FHP_WAR_WAVES_PER_EU;

// Hit log of grid thread gtid (of G): a.logDepth entries [entry][thread] in a.hitLog, which holds
// at least G * logDepth words; none without a log.
FHP_DEVI HitLog warHitLog(const WarArgs& a, uint32_t gtid, uint32_t G) {
    if (!a.hitLog || a.logDepth == 0) return noLog();
    return HitLog{(GlobalU32*)(a.hitLog + gtid), G, a.logDepth};
}

template <int B, int D, bool ADJ, bool FROZEN, bool TRANS = false>
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(kWarWavesPerEU))) void kWar(SceneDev s, WarArgs a,
                                                                                                 TraceScratch ts) {
    __shared__ uint32_t lds[D * B];
    TraceStack st = makeStack<B, D>(lds, ts, blockIdx.x, threadIdx.x);
    const uint32_t n = a.W * a.H * a.spp * (a.views ? a.nViews : 1u);
    const uint32_t G = gridDim.x * B;
    unsigned long long nc = 0, nv = 0;
    const HitLog hl = warHitLog(a, blockIdx.x * B + threadIdx.x, G);
    for (uint32_t id = blockIdx.x * B + threadIdx.x; id < n; id += G) warItem<ADJ, FROZEN, TRANS>(s, a, id, st, nc, nv, hl);
    if (a.rays) {
        atomicAdd(&a.rays[0], nc);
        atomicAdd(&a.rays[1], nv);
    }
}

template <int B, int D, bool FROZEN, bool TRANS>
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(kWarWavesPerEU))) void kWarDebug(SceneDev s, WarArgs a,
                                                                                                      TraceScratch ts) {
    __shared__ uint32_t lds[D * B];
    TraceStack st = makeStack<B, D>(lds, ts, blockIdx.x, threadIdx.x);
    const uint32_t n = a.W * a.H * a.spp * (a.views ? a.nViews : 1u);
    const uint32_t G = gridDim.x * B;
    unsigned long long nc = 0, nv = 0;
    const HitLog hl = warHitLog(a, blockIdx.x * B + threadIdx.x, G);
    for (uint32_t id = blockIdx.x * B + threadIdx.x; id < n; id += G) warDebugItem<FROZEN, TRANS>(s, a, id, st, nc, nv, hl);
    if (a.rays) {
        atomicAdd(&a.rays[0], nc);
        atomicAdd(&a.rays[1], nv);
    }
}

}  // namespace
}  // namespace fhp
