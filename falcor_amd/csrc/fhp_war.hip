// WARDiffPathTracer primal and adjoint kernels (one path per thread); the per-path code is in
// fhp_war_path.h.
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

template <int B, int D, bool ADJ, bool FROZEN>
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(kWarWavesPerEU))) void kWar(SceneDev s, WarArgs a,
                                                                                                 TraceScratch ts) {
    __shared__ uint32_t lds[D * B];
    TraceStack st = makeStack<B, D>(lds, ts, blockIdx.x, threadIdx.x);
    const uint32_t n = a.W * a.H * a.spp;
    const uint32_t G = gridDim.x * B;
    unsigned long long nc = 0, nv = 0;
    for (uint32_t id = blockIdx.x * B + threadIdx.x; id < n; id += G) warItem<ADJ, FROZEN>(s, a, id, st, nc, nv);
    if (a.rays) {
        atomicAdd(&a.rays[0], nc);
        atomicAdd(&a.rays[1], nv);
    }
}

template <int B, int D, bool FROZEN>
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(kWarWavesPerEU))) void kWarDebug(SceneDev s, WarArgs a,
                                                                                                      TraceScratch ts) {
    __shared__ uint32_t lds[D * B];
    TraceStack st = makeStack<B, D>(lds, ts, blockIdx.x, threadIdx.x);
    const uint32_t n = a.W * a.H * a.spp;
    const uint32_t G = gridDim.x * B;
    unsigned long long nc = 0, nv = 0;
    for (uint32_t id = blockIdx.x * B + threadIdx.x; id < n; id += G) warDebugItem<FROZEN>(s, a, id, st, nc, nv);
    if (a.rays) {
        atomicAdd(&a.rays[0], nc);
        atomicAdd(&a.rays[1], nv);
    }
}

// Per-pixel averages of the ForwardDiffDebug samples: the primal colour, and the red derivative
// of the samples within the threshold (WARDiffPathTracer.rt.slang:193-198: a sample whose
// derivative exceeds it leaves its pixel's derivative at zero), as AccumulatePass's Single mode
// averages frames.
__global__ void kWarDebugResolve(const float4* tmp, uint32_t npix, uint32_t spp, float threshold, float4* out,
                                 float4* dColor) {
    uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= npix) return;
    f3 sum = mk3(0.f);
    float dsum = 0.f;
    for (uint32_t k = 0; k < spp; k++) {
        float4 v = tmp[size_t(pix) * spp + k];
        sum = sum + mk3(v.x, v.y, v.z);
        if (v.w >= -threshold && v.w <= threshold) dsum += v.w;
    }
    const float w = 1.0f / float(spp);
    out[pix] = make_float4(sum.x * w, sum.y * w, sum.z * w, 1.f);
    if (dColor) dColor[pix] = make_float4(dsum * w, 0.f, 0.f, 1.f);
}

// Per-pixel average over samples (AccumulatePass Single mode, Accumulate.cs.slang:76-114).
__global__ void kWarResolve(const float4* tmp, uint32_t npix, uint32_t spp, float4* out) {
    uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= npix) return;
    f3 sum = mk3(0.f);
    for (uint32_t k = 0; k < spp; k++) {
        float4 v = tmp[size_t(pix) * spp + k];
        sum = sum + mk3(v.x, v.y, v.z);
    }
    const float w = 1.0f / float(spp);
    out[pix] = make_float4(sum.x * w, sum.y * w, sum.z * w, 1.f);
}

}  // namespace

hipError_t launchWarPrimal(const SceneDev& s, const WarArgs& a, const TraceScratch& ts, hipStream_t st, KernelTimer* tm) {
    const uint32_t n = a.W * a.H * a.spp;
    WarArgs b = a;
    b.out = a.tmp;
    if (tm) tm->begin(KK_WAR_PRIMAL, st);
    hipError_t e = withTraceConfig(ts, [&](auto Bc, auto Dc) {
        constexpr int B = decltype(Bc)::value, D = decltype(Dc)::value;
        const uint32_t blocks = std::min<uint32_t>(ts.gridBlocks, (n + B - 1) / B);
        if (samplingFrozen(s)) hipLaunchKernelGGL((kWar<B, D, false, true>), dim3(blocks), dim3(B), 0, st, s, b, ts);
        else hipLaunchKernelGGL((kWar<B, D, false, false>), dim3(blocks), dim3(B), 0, st, s, b, ts);
        return hipGetLastError();
    });
    if (tm) tm->end(st);
    if (e != hipSuccess) return e;
    const uint32_t npix = a.W * a.H;
    hipLaunchKernelGGL(kWarResolve, dim3((npix + 255) / 256), dim3(256), 0, st, a.tmp, npix, a.spp, a.out);
    return hipGetLastError();
}

hipError_t launchWarDebugImage(const SceneDev& s, const WarArgs& a, const TraceScratch& ts, hipStream_t st, KernelTimer* tm) {
    const uint32_t n = a.W * a.H * a.spp;
    if (tm) tm->begin(KK_WAR_ADJOINT, st);
    hipError_t e = withTraceConfig(ts, [&](auto Bc, auto Dc) {
        constexpr int B = decltype(Bc)::value, D = decltype(Dc)::value;
        const uint32_t blocks = std::min<uint32_t>(ts.gridBlocks, (n + B - 1) / B);
        if (samplingFrozen(s)) hipLaunchKernelGGL((kWarDebug<B, D, true>), dim3(blocks), dim3(B), 0, st, s, a, ts);
        else hipLaunchKernelGGL((kWarDebug<B, D, false>), dim3(blocks), dim3(B), 0, st, s, a, ts);
        return hipGetLastError();
    });
    if (tm) tm->end(st);
    if (e != hipSuccess) return e;
    const uint32_t npix = a.W * a.H;
    hipLaunchKernelGGL(kWarDebugResolve, dim3((npix + 255) / 256), dim3(256), 0, st, a.tmp, npix, a.spp, a.dbgThreshold,
                       a.out, a.dColor);
    return hipGetLastError();
}

hipError_t launchWarAdjoint(const SceneDev& s, const WarArgs& a, const TraceScratch& ts, hipStream_t st, KernelTimer* tm) {
    const uint32_t n = a.W * a.H * a.spp;
    if (tm) tm->begin(KK_WAR_ADJOINT, st);
    hipError_t e = withTraceConfig(ts, [&](auto Bc, auto Dc) {
        constexpr int B = decltype(Bc)::value, D = decltype(Dc)::value;
        const uint32_t blocks = std::min<uint32_t>(ts.gridBlocks, (n + B - 1) / B);
        if (samplingFrozen(s)) hipLaunchKernelGGL((kWar<B, D, true, true>), dim3(blocks), dim3(B), 0, st, s, a, ts);
        else hipLaunchKernelGGL((kWar<B, D, true, false>), dim3(blocks), dim3(B), 0, st, s, a, ts);
        return hipGetLastError();
    });
    if (tm) tm->end(st);
    return e;
}

}  // namespace fhp
