// Host check of the BSDF adjoints (test infrastructure, not shipped).
//
// Compiled for the CPU with FHP_HOST_CHECK (build.py). For random StandardMaterial parameter rows
// (with and without transmission, thin and not), random shading frames seen from the front and
// from the back, and random outgoing directions on both sides, the adjoint the kernels run
// (matAdjoint, the transmission-capable instance) is compared with central differences of the
// BSDF evaluation the kernels run (matEvalGated: lobe-selection probabilities held at the
// unperturbed parameters, which is what the adjoint differentiates). The gradient of each
// parameter of the 20-float layout is assembled as the kernels' gradient flush does (IoR through
// F0 and through the relative IoR). The reference derivative is the Richardson extrapolation
// (4 d(h/2) - d(h)) / 3 of central differences with steps h = 8e-3 and h/2, which cancels their
// O(h^2) error while the steps stay large against the float evaluation's rounding noise. It is
// only compared where the evaluation is smooth around the point: d(h) and d(h/2) must agree to
// 5%, and the one-sided differences of step h/2 to 10% (the evaluation has steps at the cosine
// cut-offs, the delta bound of the roughness and total internal reflection, and kinks where the
// transmission lobe's |wi.h wo.h| factor crosses zero).
//
// Usage: bsdfcheck <trials> <seed>. Prints the number of compared and skipped derivatives, the
// largest relative error, and one line per failure; exits non-zero on any failure.
#include "../../falcor_amd/csrc/fhp_war_path.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

using namespace fhp;

namespace {

struct Param {
    const char* name;
    int field;  // index in the 20-float row
};
const Param kParams[] = {{"baseColor.r", M_BASE},  {"baseColor.g", M_BASE + 1},    {"baseColor.b", M_BASE + 2},
                         {"roughness", M_ROUGH},   {"metallic", M_METAL},          {"ior", M_IOR},
                         {"transmission.r", M_TRANS}, {"transmission.g", M_TRANS + 1}, {"transmission.b", M_TRANS + 2},
                         {"specularTransmission", M_SPEC_TRANS}, {"diffuseTransmission", M_DIFF_TRANS}};

// The gradient of one parameter from a vertex's MatGrad, as flushGrad writes it.
double paramGrad(const MatGrad& g, const float* row, int field) {
    const float ior = row[M_IOR];
    const float f = (ior - 1.f) / (ior + 1.f);
    switch (field) {
        case M_BASE: return g.base.x;
        case M_BASE + 1: return g.base.y;
        case M_BASE + 2: return g.base.z;
        case M_ROUGH: return g.rough;
        case M_METAL: return g.metal;
        case M_IOR: return double(g.ior) * 2.0 * f * (2.0 / ((ior + 1.0) * (ior + 1.0))) + g.iorEta;
        case M_TRANS: return g.trans.x;
        case M_TRANS + 1: return g.trans.y;
        case M_TRANS + 2: return g.trans.z;
        case M_SPEC_TRANS: return g.specTrans;
        case M_DIFF_TRANS: return g.diffTrans;
        default: return 0.0;
    }
}

f3 unitVec(std::mt19937& rng) {
    std::normal_distribution<float> n;
    for (;;) {
        f3 v = mk3(n(rng), n(rng), n(rng));
        const float l = sqrtf(dot(v, v));
        if (l > 1e-3f) return v / l;
    }
}

struct Case {
    SceneDev s{};
    float row[M_STRIDE] = {};
    Shading sd0{};
    f3 wo;
};

// One evaluation at a parameter row: loadMaterial (frame flip of double-sided back faces) and the
// gated evaluation, channel c.
f3 evalAt(const Case& k, const float* row, const Material& mg) {
    Shading sd = k.sd0;
    Material m = loadMaterial(k.s, row, nullptr, k.sd0, sd, nullptr);
    return matEvalGated(sd, m, mg, k.wo);
}

}  // namespace

int main(int argc, char** argv) {
    const int trials = argc > 1 ? std::atoi(argv[1]) : 4000;
    const unsigned seed = argc > 2 ? unsigned(std::atoi(argv[2])) : 1u;
    std::mt19937 rng(seed);
    std::uniform_real_distribution<float> u(0.f, 1.f);
    long compared = 0, skipped = 0, failed = 0, nonzero = 0;
    constexpr int kNP = int(sizeof(kParams) / sizeof(kParams[0]));
    long nonzeroBy[2][kNP] = {};  // [transmission lobe sampled][param]: compared derivatives away from 0
    double worst = 0.0;
    for (int t = 0; t < trials; t++) {
        Case k;
        float* r = k.row;
        k.s.mats = r;  // material 0: the thin flag the adjoint reads
        r[M_TYPE] = float(MAT_STANDARD);
        for (int c = 0; c < 3; c++) r[M_BASE + c] = 0.1f + 0.8f * u(rng);
        r[M_ROUGH] = 0.2f + 0.75f * u(rng);
        r[M_METAL] = t % 5 == 0 ? 0.f : 0.8f * u(rng);
        r[M_IOR] = 1.2f + 0.8f * u(rng);
        r[M_DOUBLE_SIDED] = 1.f;
        r[M_TEX] = -1.f;
        const int kind = t % 4;  // 0: reflection only, 1: specular transmission, 2: diffuse, 3: both
        r[M_SPEC_TRANS] = (kind & 1) ? 0.1f + 0.8f * u(rng) : 0.f;
        r[M_DIFF_TRANS] = (kind & 2) ? 0.1f + 0.8f * u(rng) : 0.f;
        for (int c = 0; c < 3; c++) r[M_TRANS + c] = 0.2f + 0.7f * u(rng);
        const bool thin = (t / 4) % 2 == 1;
        r[M_THIN] = thin ? 1.f : 0.f;

        // Geometric normal along the shading normal up to a small tilt; the viewer on either side.
        Shading& sd = k.sd0;
        const f3 n = unitVec(rng);
        sd.faceN = n;
        setFrame(sd, normalize(n + unitVec(rng) * 0.2f));
        f3 v = unitVec(rng);
        sd.V = v;
        sd.frontFacing = dot(v, n) >= 0.f;
        sd.mat = 0;
        k.wo = unitVec(rng);

        Shading sdm = sd;
        const Material m0 = loadMaterial(k.s, r, nullptr, sd, sdm, nullptr);
        if (!m0.doubleSided) return 3;
        for (int ch = 0; ch < 3; ch++) {
            MatGrad g;
            zero(g);
            const f3 adj = mk3(ch == 0 ? 1.f : 0.f, ch == 1 ? 1.f : 0.f, ch == 2 ? 1.f : 0.f);
            matAdjoint<true>(k.s, sdm, m0, k.wo, adj, g);
            const float f0 = dot(evalAt(k, r, m0), adj);
            for (int pi = 0; pi < kNP; pi++) {
                const Param& p = kParams[pi];
                auto at = [&](float h) {
                    float rp[M_STRIDE];
                    for (int i = 0; i < M_STRIDE; i++) rp[i] = r[i];
                    rp[p.field] += h;
                    return double(dot(evalAt(k, rp, m0), adj));
                };
                auto fd = [&](float h) { return (at(h) - at(-h)) / (2.0 * h); };
                const double ad = paramGrad(g, r, p.field);
                const float h = 8e-3f;
                const double d1 = fd(h), dh = fd(0.5f * h);
                const double d2 = (4.0 * dh - d1) / 3.0;
                // Absolute floor for derivatives near zero and the differences' rounding noise.
                const double floor = 1e-4 * (std::fabs(double(f0)) + 1e-3);
                const double fwd = (at(0.5f * h) - double(f0)) / (0.5 * h), bwd = (double(f0) - at(-0.5f * h)) / (0.5 * h);
                if (std::fabs(d1 - dh) > 0.05 * std::fabs(dh) + floor || std::fabs(fwd - bwd) > 0.1 * std::fabs(dh) + floor) {
                    skipped++;
                    continue;
                }
                compared++;
                if (std::fabs(d2) > floor) {
                    nonzero++;
                    nonzeroBy[dot(sdm.toLocal(k.wo), sdm.toLocal(sdm.V)) < 0.f][pi]++;
                }
                const double err = std::fabs(ad - d2);
                const double rel = err / (std::fabs(d2) + floor);
                if (rel > worst) worst = rel;
                if (err > 0.02 * std::fabs(d2) + floor) {
                    failed++;
                    if (failed <= 20)
                        std::printf("FAIL trial=%d kind=%d thin=%d front=%d ch=%d %s adjoint=%.6e fd=%.6e f=%.4e\n", t, kind,
                                    int(thin), int(sd.frontFacing), ch, p.name, ad, d2, double(f0));
                }
            }
        }
    }
    for (int pi = 0; pi < kNP; pi++)
        std::printf("param %s nonzero_reflection=%ld nonzero_transmission=%ld\n", kParams[pi].name, nonzeroBy[0][pi],
                    nonzeroBy[1][pi]);
    std::printf("compared=%ld nonzero=%ld skipped=%ld failed=%ld worst_rel=%.3e\n", compared, nonzero, skipped, failed,
                worst);
    return failed ? 1 : 0;
}
