// WARDiffPathTracer primal and adjoint launches (one path per thread); the per-path code is in
// fhp_war_path.h, the kernels in fhp_war_kernels.h.
#include "fhp_war_kernels.h"

namespace fhp {

namespace {

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

// Per-pixel average over samples (AccumulatePass Single mode, Accumulate.cs.slang:76-114); with
// batched views, pixel v * npix + p of the sample array resolves into pixel p of view v's image.
__global__ void kWarResolve(const float4* tmp, uint32_t npix, uint32_t nviews, uint32_t spp, float4* out,
                            const WarView* views) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= npix * nviews) return;
    f3 sum = mk3(0.f);
    for (uint32_t k = 0; k < spp; k++) {
        float4 v = tmp[size_t(g) * spp + k];
        sum = sum + mk3(v.x, v.y, v.z);
    }
    const float w = 1.0f / float(spp);
    const uint32_t view = g / npix, pix = g - view * npix;
    float4* o = views ? views[view].out : out;
    o[pix] = make_float4(sum.x * w, sum.y * w, sum.z * w, 1.f);
}

}  // namespace

hipError_t launchWarPrimal(const SceneDev& s, const WarArgs& a, const TraceScratch& ts, hipStream_t st, KernelTimer* tm) {
    const uint32_t nviews = a.views ? a.nViews : 1u;
    const uint32_t n = a.W * a.H * a.spp * nviews;
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
    hipLaunchKernelGGL(kWarResolve, dim3((npix * nviews + 255) / 256), dim3(256), 0, st, a.tmp, npix, nviews, a.spp, a.out,
                       a.views);
    return hipGetLastError();
}

hipError_t launchWarDebugImage(const SceneDev& s, const WarArgs& a, const TraceScratch& ts, hipStream_t st, KernelTimer* tm) {
    const uint32_t n = a.W * a.H * a.spp;
    if (tm) tm->begin(KK_WAR_ADJOINT, st);
    hipError_t e = withTraceConfig(ts, [&](auto Bc, auto Dc) {
        constexpr int B = decltype(Bc)::value, D = decltype(Dc)::value;
        const uint32_t blocks = std::min<uint32_t>(ts.gridBlocks, (n + B - 1) / B);
        if (s.hasTransmission) return launchWarTransKernel(s, a, ts, st, blocks, true);
        if (samplingFrozen(s)) hipLaunchKernelGGL((kWarDebug<B, D, true, false>), dim3(blocks), dim3(B), 0, st, s, a, ts);
        else hipLaunchKernelGGL((kWarDebug<B, D, false, false>), dim3(blocks), dim3(B), 0, st, s, a, ts);
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
    const uint32_t n = a.W * a.H * a.spp * (a.views ? a.nViews : 1u);
    if (tm) tm->begin(KK_WAR_ADJOINT, st);
    hipError_t e = withTraceConfig(ts, [&](auto Bc, auto Dc) {
        constexpr int B = decltype(Bc)::value, D = decltype(Dc)::value;
        const uint32_t blocks = std::min<uint32_t>(ts.gridBlocks, (n + B - 1) / B);
        // Scenes with transmissive materials use the instances that carry the transmission adjoint.
        if (s.hasTransmission) return launchWarTransKernel(s, a, ts, st, blocks, false);
        if (samplingFrozen(s)) hipLaunchKernelGGL((kWar<B, D, true, true, false>), dim3(blocks), dim3(B), 0, st, s, a, ts);
        else hipLaunchKernelGGL((kWar<B, D, true, false, false>), dim3(blocks), dim3(B), 0, st, s, a, ts);
        return hipGetLastError();
    });
    if (tm) tm->end(st);
    return e;
}

}  // namespace fhp
