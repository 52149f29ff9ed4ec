// WARDiffPathTracer adjoint and ForwardDiffDebug kernels with the transmission-lobe adjoint
// (scenes with transmissive StandardMaterials); launched from fhp_war.hip.
#include "fhp_war_kernels.h"

namespace fhp {

hipError_t launchWarTransKernel(const SceneDev& s, const WarArgs& a, const TraceScratch& ts, hipStream_t st,
                                uint32_t blocks, bool debug) {
    return withTraceConfig(ts, [&](auto Bc, auto Dc) {
        constexpr int B = decltype(Bc)::value, D = decltype(Dc)::value;
        const bool fz = samplingFrozen(s);
        if (debug && fz) hipLaunchKernelGGL((kWarDebug<B, D, true, true>), dim3(blocks), dim3(B), 0, st, s, a, ts);
        else if (debug) hipLaunchKernelGGL((kWarDebug<B, D, false, true>), dim3(blocks), dim3(B), 0, st, s, a, ts);
        else if (fz) hipLaunchKernelGGL((kWar<B, D, true, true, true>), dim3(blocks), dim3(B), 0, st, s, a, ts);
        else hipLaunchKernelGGL((kWar<B, D, true, false, true>), dim3(blocks), dim3(B), 0, st, s, a, ts);
        return hipGetLastError();
    });
}

}  // namespace fhp
