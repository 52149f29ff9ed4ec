// Frame accumulation and image comparison kernels: the output side of the path tracer.
//
// Both are streaming kernels bound by HBM bandwidth: one lane per pixel, 16-byte vector loads
// and stores of RGBA rows, no reuse. They follow the written operation order exactly
// (-ffp-contract=off for the whole library), so results match a sequential CPU restatement
// bit for bit, except the total of an image comparison, which is a sum over pixels in a
// different (fixed) order.
#include "fhp_internal.h"

namespace fhp {
namespace {

constexpr int kImgBlock = 256;

template <int FMT> __device__ __forceinline__ float4 loadInput(const void* in, uint64_t i) {
    if constexpr (FMT == kAccumInputFloat) {
        return reinterpret_cast<const float4*>(in)[i];
    } else if constexpr (FMT == kAccumInputUint) {
        const uint4 u = reinterpret_cast<const uint4*>(in)[i];
        return make_float4(float(u.x), float(u.y), float(u.z), float(u.w));
    } else {
        const int4 s = reinterpret_cast<const int4*>(in)[i];
        return make_float4(float(s.x), float(s.y), float(s.z), float(s.w));
    }
}

__device__ __forceinline__ float4 lerp4(float4 x, float4 y, float s) {
    // lerp(x, y, s) = x * (1 - s) + y * s, the form the double-precision path writes out
    // (Accumulate.cs.slang:92, :181).
    const float r = 1.f - s;
    return make_float4(x.x * r + y.x * s, x.y * r + y.y * s, x.z * r + y.z * s, x.w * r + y.w * s);
}

// One frame of AccumulatePass over n pixels (Accumulate.cs.slang:76-190).
// MODE: kAccumDouble (sum as double4 in sumD), kAccumSingle (float4 sum),
// kAccumSingleCompensated (float4 sum + Kahan correction). count = frames accumulated before
// this one (gAccumCount). flags: kAccumEnabled (gAccumulate), kAccumMovingAverage
// (gMovingAverageMode), kAccumFirst (state is treated as cleared: the reference clears the
// buffers when the frame count is 0, AccumulatePass.cpp:401-407).
template <int MODE, int FMT>
__global__ __launch_bounds__(kImgBlock) void kAccumulate(const void* __restrict__ in, float4* __restrict__ out,
                                                         float4* __restrict__ sum, float4* __restrict__ corr,
                                                         double2* __restrict__ sumD, uint64_t n, uint32_t count,
                                                         uint32_t flags) {
    const bool enabled = flags & kAccumEnabled;
    const bool mavg = flags & kAccumMovingAverage;
    const bool first = flags & kAccumFirst;
    const uint64_t G = uint64_t(gridDim.x) * kImgBlock;
    for (uint64_t i = uint64_t(blockIdx.x) * kImgBlock + threadIdx.x; i < n; i += G) {
        const float4 cur = loadInput<FMT>(in, i);
        if (!enabled) {
            out[i] = cur;
            continue;
        }
        if constexpr (MODE == kAccumSingle) {
            const float w = 1.f / float(count + 1u);
            const float4 prev = first ? make_float4(0.f, 0.f, 0.f, 0.f) : sum[i];
            float4 o;
            if (mavg) {
                o = lerp4(prev, cur, w);
                sum[i] = o;
            } else {
                const float4 s = make_float4(prev.x + cur.x, prev.y + cur.y, prev.z + cur.z, prev.w + cur.w);
                o = make_float4(s.x * w, s.y * w, s.z * w, s.w * w);
                sum[i] = s;
            }
            out[i] = o;
        } else if constexpr (MODE == kAccumSingleCompensated) {
            const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
            const float4 s = first ? z : sum[i];
            const float4 c = first ? z : corr[i];
            const float4 y = make_float4(cur.x - c.x, cur.y - c.y, cur.z - c.z, cur.w - c.w);
            const float4 sn = make_float4(s.x + y.x, s.y + y.y, s.z + y.z, s.w + y.w);
            const float d = float(count + 1u);
            out[i] = make_float4(sn.x / d, sn.y / d, sn.z / d, sn.w / d);
            sum[i] = sn;
            corr[i] = make_float4((sn.x - s.x) - y.x, (sn.y - s.y) - y.y, (sn.z - s.z) - y.z, (sn.w - s.w) - y.w);
        } else {
            const double w = 1.0 / double(count + 1u);
            const double2 z = make_double2(0.0, 0.0);
            double2 lo = first ? z : sumD[2 * i];
            double2 hi = first ? z : sumD[2 * i + 1];
            double s[4] = {lo.x, lo.y, hi.x, hi.y};
            const float c[4] = {cur.x, cur.y, cur.z, cur.w};
            float o[4];
            for (int k = 0; k < 4; k++) {
                if (mavg) {
                    s[k] = s[k] * (1.0 - w) + double(c[k]) * w;
                    o[k] = float(s[k]);
                } else {
                    s[k] += double(c[k]);
                    o[k] = float(s[k] * w);
                }
            }
            sumD[2 * i] = make_double2(s[0], s[1]);
            sumD[2 * i + 1] = make_double2(s[2], s[3]);
            out[i] = make_float4(o[0], o[1], o[2], o[3]);
        }
    }
}

template <int MODE>
hipError_t launchAccumulateMode(const AccumArgs& a, uint32_t blocks, hipStream_t st) {
    switch (a.inputFormat) {
    case kAccumInputFloat:
        hipLaunchKernelGGL((kAccumulate<MODE, kAccumInputFloat>), dim3(blocks), dim3(kImgBlock), 0, st, a.in, a.out, a.sum,
                           a.corr, a.sumD, a.n, a.count, a.flags);
        break;
    case kAccumInputUint:
        hipLaunchKernelGGL((kAccumulate<MODE, kAccumInputUint>), dim3(blocks), dim3(kImgBlock), 0, st, a.in, a.out, a.sum,
                           a.corr, a.sumD, a.n, a.count, a.flags);
        break;
    default:
        hipLaunchKernelGGL((kAccumulate<MODE, kAccumInputSint>), dim3(blocks), dim3(kImgBlock), 0, st, a.in, a.out, a.sum,
                           a.corr, a.sumD, a.n, a.count, a.flags);
        break;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------- image comparison
// Per-pixel error over the first nch channels (ImageCompare.cpp:191-233): the channel terms
// are formed in float where the reference forms them in float and summed in double.
template <int METRIC> __device__ __forceinline__ double pixelError(float4 a4, float4 b4, int nch) {
    const float a[4] = {a4.x, a4.y, a4.z, a4.w};
    const float b[4] = {b4.x, b4.y, b4.z, b4.w};
    double e = 0.0;
    for (int k = 0; k < nch; k++) {
        const float d = a[k] - b[k];
        if constexpr (METRIC == kMetricMSE) {
            e += double(d * d);
        } else if constexpr (METRIC == kMetricRMSE) {
            e += double(d * d) / (double(a[k] * a[k]) + 1e-3);
        } else if constexpr (METRIC == kMetricMAE) {
            e += double(fabsf(d * d));  // the reference's MAE squares the difference too (:213-222)
        } else {
            e += fabs(double(d) / (double(a[k]) + 1e-3));
        }
    }
    if constexpr (METRIC == kMetricMAPE) return 100.0 * e / double(nch);
    return e / double(nch);
}

template <typename T> __device__ __forceinline__ T waveSum(T v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Pass 1: per-pixel errors (written to errMap when given) and one partial per block: the sum of
// the block's pixel errors (pixels in grid-stride order, lanes in a fixed tree) and the min and
// max of the float errors, NaN errors left out.
template <int METRIC>
__global__ __launch_bounds__(kImgBlock) void kCompare(const float4* __restrict__ a, const float4* __restrict__ b, uint64_t n,
                                                      int nch, float* __restrict__ errMap, CompareResult* __restrict__ part) {
    __shared__ double sSum[kImgBlock / 64];
    __shared__ float sMin[kImgBlock / 64], sMax[kImgBlock / 64];
    double acc = 0.0;
    float mn = INFINITY, mx = -INFINITY;
    const uint64_t G = uint64_t(gridDim.x) * kImgBlock;
    for (uint64_t i = uint64_t(blockIdx.x) * kImgBlock + threadIdx.x; i < n; i += G) {
        const double e = pixelError<METRIC>(a[i], b[i], nch);
        const float ef = float(e);
        if (errMap) errMap[i] = ef;
        acc += e;
        if (ef == ef) {
            mn = fminf(mn, ef);
            mx = fmaxf(mx, ef);
        }
    }
    acc = waveSum(acc);
    for (int o = 32; o > 0; o >>= 1) {
        mn = fminf(mn, __shfl_down(mn, o, 64));
        mx = fmaxf(mx, __shfl_down(mx, o, 64));
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        sSum[wave] = acc;
        sMin[wave] = mn;
        sMax[wave] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        CompareResult r{0.0, INFINITY, -INFINITY};
        for (int w = 0; w < kImgBlock / 64; w++) {
            r.sum += sSum[w];
            r.minErr = fminf(r.minErr, sMin[w]);
            r.maxErr = fmaxf(r.maxErr, sMax[w]);
        }
        part[blockIdx.x] = r;
    }
}

// Pass 2 (one block): partials in block order -> result.
__global__ __launch_bounds__(kImgBlock) void kCompareFinal(const CompareResult* __restrict__ part, uint32_t nPart,
                                                           CompareResult* __restrict__ res) {
    __shared__ double sSum[kImgBlock / 64];
    __shared__ float sMin[kImgBlock / 64], sMax[kImgBlock / 64];
    double acc = 0.0;
    float mn = INFINITY, mx = -INFINITY;
    for (uint32_t i = threadIdx.x; i < nPart; i += kImgBlock) {
        const CompareResult p = part[i];
        acc += p.sum;
        mn = fminf(mn, p.minErr);
        mx = fmaxf(mx, p.maxErr);
    }
    acc = waveSum(acc);
    for (int o = 32; o > 0; o >>= 1) {
        mn = fminf(mn, __shfl_down(mn, o, 64));
        mx = fmaxf(mx, __shfl_down(mx, o, 64));
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        sSum[wave] = acc;
        sMin[wave] = mn;
        sMax[wave] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        CompareResult r{0.0, INFINITY, -INFINITY};
        for (int w = 0; w < kImgBlock / 64; w++) {
            r.sum += sSum[w];
            r.minErr = fminf(r.minErr, sMin[w]);
            r.maxErr = fmaxf(r.maxErr, sMax[w]);
        }
        *res = r;
    }
}

// Heat-map ramp stops (r, g, b): blue, teal, green, yellow, red.
__constant__ float kRamp[5][3] = {{0.f, 0.f, 1.f}, {0.f, 1.f, 1.f}, {0.f, 1.f, 0.f}, {1.f, 1.f, 0.f}, {1.f, 0.f, 0.f}};

// Heat map of an error map (ImageCompare.cpp:242-270): t = clamp((e - min) / max(1e-5, max - min)),
// colour ramp blue -> teal -> green -> yellow -> red in four linear segments, alpha 1.
__global__ __launch_bounds__(kImgBlock) void kHeatMap(const float* __restrict__ err, uint64_t n,
                                                      const CompareResult* __restrict__ res, float4* __restrict__ out) {
    const float mn = res->minErr, mx = res->maxErr;
    const float d = mx - mn;
    const float range = (1e-5f < d) ? d : 1e-5f;  // std::max(1e-5f, d)
    const uint64_t G = uint64_t(gridDim.x) * kImgBlock;
    for (uint64_t i = uint64_t(blockIdx.x) * kImgBlock + threadIdx.x; i < n; i += G) {
        const float x = (err[i] - mn) / range;
        // clamp(x, 0, 1) = std::max(0, std::min(1, x)): a NaN x gives 1.
        const float m = (x < 1.f) ? x : 1.f;
        const float t = (0.f < m) ? m : 0.f;
        int c = int(floorf(t * 4.f));
        c = c < 0 ? 0 : (c > 3 ? 3 : c);
        const float u = t * 4.f - float(c);
        const float* c0 = kRamp[c];
        const float* c1 = kRamp[c + 1];
        out[i] = make_float4(c0[0] + u * (c1[0] - c0[0]), c0[1] + u * (c1[1] - c0[1]), c0[2] + u * (c1[2] - c0[2]), 1.f);
    }
}

uint32_t streamBlocks(uint64_t n, uint32_t cap) {
    return uint32_t(std::min<uint64_t>(std::max<uint64_t>((n + kImgBlock - 1) / kImgBlock, 1), cap));
}

}  // namespace

hipError_t launchAccumulate(const AccumArgs& a, hipStream_t st) {
    if (a.n == 0) return hipSuccess;
    const uint32_t blocks = streamBlocks(a.n, 1u << 20);
    switch (a.mode) {
    case kAccumDouble: return launchAccumulateMode<kAccumDouble>(a, blocks, st);
    case kAccumSingle: return launchAccumulateMode<kAccumSingle>(a, blocks, st);
    default: return launchAccumulateMode<kAccumSingleCompensated>(a, blocks, st);
    }
}

uint32_t comparePartials(uint64_t n) { return streamBlocks(n, kComparePartialsMax); }

hipError_t launchCompare(int metric, const float4* a, const float4* b, uint64_t n, int nch, float* errMap,
                         CompareResult* part, CompareResult* res, hipStream_t st) {
    const uint32_t blocks = comparePartials(n);
    switch (metric) {
    case kMetricMSE:
        hipLaunchKernelGGL(kCompare<kMetricMSE>, dim3(blocks), dim3(kImgBlock), 0, st, a, b, n, nch, errMap, part);
        break;
    case kMetricRMSE:
        hipLaunchKernelGGL(kCompare<kMetricRMSE>, dim3(blocks), dim3(kImgBlock), 0, st, a, b, n, nch, errMap, part);
        break;
    case kMetricMAE:
        hipLaunchKernelGGL(kCompare<kMetricMAE>, dim3(blocks), dim3(kImgBlock), 0, st, a, b, n, nch, errMap, part);
        break;
    default:
        hipLaunchKernelGGL(kCompare<kMetricMAPE>, dim3(blocks), dim3(kImgBlock), 0, st, a, b, n, nch, errMap, part);
        break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kCompareFinal, dim3(1), dim3(kImgBlock), 0, st, part, blocks, res);
    return hipGetLastError();
}

hipError_t launchHeatMap(const float* err, uint64_t n, const CompareResult* res, float4* out, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(kHeatMap, dim3(streamBlocks(n, 1u << 20)), dim3(kImgBlock), 0, st, err, n, res, out);
    return hipGetLastError();
}

}  // namespace fhp
