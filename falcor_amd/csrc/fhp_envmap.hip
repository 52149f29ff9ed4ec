// Hierarchical importance map of the environment map (EnvMapSampler::createImportanceMap,
// EnvMapSampler.cpp:83-115, and EnvMapSamplerSetup.cs.slang:44-68). The finest level is a square
// equal-area octahedral map: each texel averages the luminance of the lat-long map over a grid of
// samples x samples points of its octahedral footprint (bilinear lookups, no intensity, tint or
// rotation: the map lives in the env map's local frame). Coarser levels are 2 x 2 means down to
// 1 x 1, which then holds the mean of the finest level. Levels are stored coarse to fine, the
// r x r level at offset (r * r - 1) / 3.
#include "fhp_internal.h"

namespace fhp {
namespace {

constexpr int kEnvBlock = 256;

__global__ void kEnvImportance(SceneDev s, float* __restrict__ out, uint32_t dim, uint32_t ns) {
    const uint32_t i = blockIdx.x * kEnvBlock + threadIdx.x;
    if (i >= dim * dim) return;
    const uint32_t x = i % dim, y = i / dim;
    const float invDimS = 1.f / float(dim * ns);
    float L = 0.f;
    for (uint32_t sy = 0; sy < ns; sy++) {
        for (uint32_t sx = 0; sx < ns; sx++) {
            const float px = (float(x * ns + sx) + 0.5f) * invDimS, py = (float(y * ns + sy) + 0.5f) * invDimS;
            const f3 d = octToDirEqualArea(px, py);
            float u, v;
            latlongUV(normalize(d), u, v);
            L += luminance(envLookup(s, u, v));
        }
    }
    out[i] = L * (1.f / float(ns * ns));
}

// One coarser level: dst (r x r) from src (2r x 2r).
__global__ void kEnvImportanceDown(const float* __restrict__ src, float* __restrict__ dst, uint32_t r) {
    const uint32_t i = blockIdx.x * kEnvBlock + threadIdx.x;
    if (i >= r * r) return;
    const uint32_t x = i % r, y = i / r, R = 2u * r;
    const float* a = src + size_t(2u * y) * R + 2u * x;
    dst[i] = ((a[0] + a[1]) + (a[R] + a[R + 1])) * 0.25f;
}

// EnvMapSampler::sample for a batch of (ux, uy): out 8 floats per sample (dir, pdf, Le, valid).
__global__ void kEnvSampleBatch(SceneDev s, const float* __restrict__ u, uint32_t n, float* __restrict__ out) {
    const uint32_t i = blockIdx.x * kEnvBlock + threadIdx.x;
    if (i >= n) return;
    f3 dir = mk3(0.f), Le = mk3(0.f);
    float pdf = 0.f;
    const bool ok = envSample(s, u[2 * size_t(i)], u[2 * size_t(i) + 1], dir, pdf, Le);
    float* o = out + 8 * size_t(i);
    o[0] = dir.x; o[1] = dir.y; o[2] = dir.z; o[3] = pdf;
    o[4] = Le.x; o[5] = Le.y; o[6] = Le.z; o[7] = ok ? 1.f : 0.f;
}

__global__ void kEnvPdfBatch(SceneDev s, const float* __restrict__ dirs, uint32_t n, float* __restrict__ pdf) {
    const uint32_t i = blockIdx.x * kEnvBlock + threadIdx.x;
    if (i >= n) return;
    pdf[i] = envPdf(s, mk3(dirs[3 * size_t(i)], dirs[3 * size_t(i) + 1], dirs[3 * size_t(i) + 2]));
}

}  // namespace

hipError_t launchEnvSampleBatch(const SceneDev& s, const float* u, uint32_t n, float* out, hipStream_t st) {
    hipLaunchKernelGGL(kEnvSampleBatch, dim3((n + kEnvBlock - 1) / kEnvBlock), dim3(kEnvBlock), 0, st, s, u, n, out);
    return hipGetLastError();
}

hipError_t launchEnvPdfBatch(const SceneDev& s, const float* dirs, uint32_t n, float* pdf, hipStream_t st) {
    hipLaunchKernelGGL(kEnvPdfBatch, dim3((n + kEnvBlock - 1) / kEnvBlock), dim3(kEnvBlock), 0, st, s, dirs, n, pdf);
    return hipGetLastError();
}

hipError_t launchEnvImportance(const SceneDev& s, float* out, uint32_t dim, uint32_t samplesPerAxis, hipStream_t st) {
    float* fine = out + (dim * dim - 1u) / 3u;
    hipLaunchKernelGGL(kEnvImportance, dim3((dim * dim + kEnvBlock - 1) / kEnvBlock), dim3(kEnvBlock), 0, st, s, fine, dim,
                       samplesPerAxis);
    for (uint32_t r = dim / 2; r >= 1; r /= 2) {
        const float* src = out + (4u * r * r - 1u) / 3u;
        float* dst = out + (r * r - 1u) / 3u;
        hipLaunchKernelGGL(kEnvImportanceDown, dim3((r * r + kEnvBlock - 1) / kEnvBlock), dim3(kEnvBlock), 0, st, src, dst, r);
    }
    return hipGetLastError();
}

}  // namespace fhp
