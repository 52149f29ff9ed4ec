// WARDiffPathTracer vertex gradients: the oracle's restatement of the estimator (test
// infrastructure only; the product never links this file).
//
// The reference differentiates its warped-area estimator (WARDiffPathTracer.rt.slang:66-143,
// PTUtils.slang:136-341, WarpedAreaReparam.slang:50-332, DiffSceneQuery.slang:131-261,
// GeometryHelpers.slang:51-73) with Slang's automatic differentiation. Every sampled decision is
// detached: the triangles the rays hit, the light sample, the BSDF sample and its pdf, the
// auxiliary directions and the triangles and barycentrics they hit, the harmonic weights, the MIS
// weights. What is differentiated is the path's radiance as a smooth function of the vertex
// positions and normals with those decisions held.
//
// This file restates that function. For a perturbation field (dp_v, dn_v) over the scene's
// vertices, F(theta) is the radiance of a path with positions p_v + theta dp_v and normals
// n_v + theta dn_v, in double precision, with the decisions taken once at theta = 0 in the
// reference's FP32 arithmetic (the oracle's WAR primal path). The warps follow the reference's
// construction: the warp is V(theta) = u(theta) - u(0) with u the normalized harmonic-weighted
// mean of the auxiliary hit directions, and the Jacobian determinant is det(I + J(theta) - J(0)),
// J being the forward derivative of u with respect to the ray direction through the harmonic
// weights' custom derivative (fwd_computeHarmonicWeight). The screen warp projects u to screen
// space (computeCameraRayScreenPos). The derivative sum_paths dLdI . dF/dtheta at 0, taken by
// central differences of the smooth function F, is the directional derivative the adjoint's
// vertex gradients project to: sum_v (dS/dp_v . dp_v + dS/dn_v . dn_v).
#include "fo_scene.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <limits>
#include <thread>
#include <vector>

namespace fo {
namespace {

using V3d = V3<double>;

V3d widen(V3f v) { return V3d(double(v.x), double(v.y), double(v.z)); }
V3d normalizeSafeD(V3d v) {
    const double l = length(v);
    return l > 0.0 ? v / l : V3d(0.0);
}

// One auxiliary ray of traceAsymptoticWeightedMeanIntersection (WarpedAreaReparam.slang:111-193):
// direction, hit (AttachToGeometry: triangle and barycentrics held), harmonic weight w and the
// coefficient c of its custom direction derivative, dW = c * dot(-dir x (dir x d), a).
struct Aux {
    V3d a;
    bool hit;
    uint32_t tri;
    double b[3];
    double w, c;
};
struct Warp {
    bool used = false;   // reparameterized (useWAR, the ray exists)
    bool valid = false;  // total weight > 0: otherwise the warp returns the input directions
    V3d origin0;         // detached ray origin (FP32 value)
    std::vector<Aux> aux;
};

// Auxiliary rays from a copy of the path's sample generator (FP32, as the reference computes them).
void traceAux(const Scene& s, Tracer& tr, SampleGen sg, V3f origin, V3f baseDir, float kappa, uint32_t n, bool anti,
              Warp& w) {
    w.used = true;
    w.origin0 = widen(origin);
    w.aux.resize(n);
    std::vector<float> rnd(n);
    for (uint32_t i = 0; i < n; i++) rnd[i] = sg.next1D();
    // ShadingFrame(baseDirection, float4(perp_stark(baseDirection), 1)): N, then T orthonormalized.
    Frame<float> f;
    f.N = normalize(baseDir);
    V3f t = perp_stark(baseDir);
    f.T_ = normalize(t - f.N * dot(t, f.N));
    f.B = cross(f.N, f.T_);
    double total = 0.0;
    for (uint32_t i = 0; i < n; i++) {
        float ux, uy;
        if (anti) {
            ux = rnd[(i / 2) * 2];
            uy = rnd[(i / 2) * 2 + 1];
        } else {
            ux = sg.next1D();
            uy = sg.next1D();
        }
        V3f dl = sampleVonMisesFisher(ux, uy, kappa);
        V3f a = normalize(f.fromLocal(dl));
        if (anti && i % 2 == 1) {
            // reflect(-a, base) = -a - 2 dot(-a, base) base
            const V3f m = -a;
            a = m - baseDir * (2.f * dot(m, baseDir));
        }
        Aux& x = w.aux[i];
        x.a = widen(a);
        Hit h = tr.closest(origin, a, 0.f, std::numeric_limits<float>::max());
        float boundary = 0.f;
        x.hit = h.valid();
        x.tri = h.tri;
        if (x.hit) {
            x.b[0] = 1.0 - double(h.u) - double(h.v);
            x.b[1] = h.u;
            x.b[2] = h.v;
            // computeShadingFrame: the interpolated vertex normal (not normalized).
            const float b0 = 1.f - h.u - h.v;
            const V3f nrm = s.nrm[s.idx[3 * h.tri]] * b0 + s.nrm[s.idx[3 * h.tri + 1]] * h.u + s.nrm[s.idx[3 * h.tri + 2]] * h.v;
            const float dn = dot(nrm, a);
            boundary = dn * dn;
        }
        // computeHarmonicWeight / fwd_computeHarmonicWeight (WarpedAreaReparam.slang:50-98)
        const float sy = std::max(1.f - uy, 1e-6f);
        const float invVMF = 1.f / ((1.f - sy) * std::exp(-2.f * kappa) + sy);
        const float wDen = invVMF - 1.f + boundary;
        const float rcp = wDen > 1e-4f ? 1.f / wDen : 0.f;
        const float wt = rcp * rcp * rcp * invVMF;
        x.w = wt;
        x.c = double(invVMF * wt * rcp * kappa * 3.f);
        total += double(wt);
    }
    w.valid = total > 0.0;
}

// The perturbed scene: positions and normals at theta.
struct Perturbed {
    const Scene& s;
    const float* dp;  // [nV][3] or null
    const float* dn;  // [nV][3] or null
    double theta;
    V3d pos(uint32_t v) const {
        V3d p = widen(s.pos[v]);
        if (dp) p = p + V3d(dp[3 * v], dp[3 * v + 1], dp[3 * v + 2]) * theta;
        return p;
    }
    V3d nrm(uint32_t v) const {
        V3d n = widen(s.nrm[v]);
        if (dn) n = n + V3d(dn[3 * v], dn[3 * v + 1], dn[3 * v + 2]) * theta;
        return n;
    }
    void tri(uint32_t t, V3d v[3]) const {
        for (int k = 0; k < 3; k++) v[k] = pos(s.idx[3 * t + k]);
    }
};

// Weighted mean direction m, normalized u, and the direction derivatives du/dd_k of u for the
// tangents d_k (through the harmonic weights' custom derivative at primal direction dir).
struct WarpEval {
    V3d u;
    V3d du[3];
};
WarpEval evalWarp(const Perturbed& P, const Warp& w, V3d origin, V3d dir, const V3d* tangents, int nt) {
    WarpEval r;
    double sw = 0.0;
    V3d sd(0.0);
    std::vector<V3d> dirW(w.aux.size());
    for (size_t i = 0; i < w.aux.size(); i++) {
        const Aux& x = w.aux[i];
        V3d d = x.a;
        if (x.hit) {
            V3d v[3];
            P.tri(x.tri, v);
            const V3d p = v[0] * x.b[0] + v[1] * x.b[1] + v[2] * x.b[2];
            d = normalizeSafeD(p - origin);
        }
        dirW[i] = d;
        sw += x.w;
        sd = sd + d * x.w;
    }
    const double swc = std::max(sw, 1e-8);
    const V3d m = sd / swc;
    const double ml = length(m);
    r.u = ml > 0.0 ? m / ml : V3d(0.0);
    for (int k = 0; k < nt; k++) {
        // dW_i = clamp(c_i dot(-dir x (dir x d_k), a_i), -1e10, 1e10)
        const V3d t2 = -cross(dir, cross(dir, tangents[k]));
        double sdw = 0.0;
        V3d sdd(0.0);
        for (size_t i = 0; i < w.aux.size(); i++) {
            const double dw = std::min(std::max(w.aux[i].c * dot(t2, w.aux[i].a), -1e10), 1e10);
            sdw += dw;
            sdd = sdd + dirW[i] * dw;
        }
        // m = sd / max(sw, 1e-8): the max passes the derivative of sw only when sw > 1e-8.
        const V3d dm = sw > 1e-8 ? (sdd - m * sdw) / swc : sdd / swc;
        r.du[k] = ml > 0.0 ? (dm - r.u * dot(r.u, dm)) / ml : V3d(0.0);
    }
    return r;
}

struct CamD {
    V3d pos, U, V, W;
};
// computeCameraRayDirection (DiffSceneQuery.slang:252-261), unnormalized, and its screen gradient.
V3d camRaw(const CamD& c, double sx, double sy, uint32_t W, uint32_t H) {
    const double px = sx / double(W), py = sy / double(H);
    return c.U * (2.0 * px - 1.0) + c.V * (-2.0 * py + 1.0) + c.W;
}
// computeCameraRayScreenPos (DiffSceneQuery.slang:221-250) and its derivative along v.
void camScreen(const CamD& c, V3d dir, V3d v, uint32_t W, uint32_t H, double& sx, double& sy, double& dsx, double& dsy) {
    const double l = length(dir);
    const V3d n = l > 0.0 ? dir / l : V3d(0.0);
    const V3d dn = l > 0.0 ? (v - n * dot(n, v)) / l : V3d(0.0);
    const double WW = dot(c.W, c.W), Wn = dot(c.W, n), Wdn = dot(c.W, dn);
    const V3d fp = n * (WW / Wn);
    const V3d dfp = dn * (WW / Wn) - n * (WW * Wdn / (Wn * Wn));
    const V3d nc = fp - c.W;
    const double UU = dot(c.U, c.U), VV = dot(c.V, c.V);
    const double ndcx = dot(nc, c.U) / UU, ndcy = dot(nc, c.V) / VV;
    const double dndcx = dot(dfp, c.U) / UU, dndcy = dot(dfp, c.V) / VV;
    sx = (ndcx + 1.0) / 2.0 * double(W) - 0.5;
    sy = (ndcy - 1.0) / -2.0 * double(H) - 0.5;
    dsx = dndcx / 2.0 * double(W);
    dsy = dndcy / -2.0 * double(H);
}

double det3(const V3d c[3]) { return dot(c[0], cross(c[1], c[2])); }

// One path vertex of F(theta), taken at theta = 0: the triangle hit by the incoming ray, the
// material set-up, the emission the vertex adds (with its held MIS weight), its NEE sample and
// warp, and its scatter sample (with the outgoing ray's FP32 origin) and warp.
struct VertexRec {
    uint32_t tri = kInvalid;
    uint32_t mat = 0;
    bool front = true, flip = false;
    V3d faceN;
    bool emit = false;
    V3d E;
    double misE = 1.0;
    bool lit = false;
    V3d lightPos, lightN, Li;
    double misN = 1.0;
    Warp nee;
    bool scat = false;
    V3d woS, oS0;
    double pdfS = 0.0;
    Warp sc;
};

// Everything a path's F(theta) needs, taken at theta = 0.
struct PathRec {
    double psx, psy, pcx, pcy;
    Warp screen;
    std::vector<VertexRec> v;
    V3d dL;
};

struct Params {
    uint32_t W, H, aux;
    float kappa, kappaS;
    bool useWAR, bsdfSampling, nee, mis, anti;
    bool neeHeld;  // extension: the NEE direction held (a solid-angle sample), no area factor
    uint32_t maxBounces;
};

// The oracle's WAR primal path at theta = 0 (fo_impl.cpp warTracePaths, the reference's tracePath
// with handleHit, PTUtils.slang:255-341), recording the decisions of every vertex and tracing the
// auxiliary rays of each warp from a copy of the sample generator at the warp (the screen warp
// before the primary ray, the NEE warp after the light sample, the scatter warp after the BSDF
// sample). Returns false for a path with no primary hit.
bool recordPath(const Scene& s, const Params& P, uint32_t px, uint32_t py, uint32_t frameSeed, Tracer& tr, PathRec& r) {
    SampleGen sg = makeSampleGen(px, py, frameSeed / 2);
    for (int i = 0; i < 6; i++) sg.next1D();
    const float pfx = float(px) + 0.5f, pfy = float(py) + 0.5f;
    const float ux = sg.next1D(), uy = sg.next1D();
    float r0 = std::sqrt(-2.f * std::log(ux));
    float jx = (r0 * std::cos(Consts<float>::two_pi * uy)) * 0.5f;
    float jy = (r0 * std::sin(Consts<float>::two_pi * uy)) * 0.5f;
    if (P.anti && frameSeed % 2 == 0) { jx = -jx; jy = -jy; }
    const float sx = pfx + jx, sy = pfy + jy;
    r.psx = sx; r.psy = sy; r.pcx = pfx; r.pcy = pfy;
    float p_x = sx / float(P.W), p_y = sy / float(P.H);
    V3f dir = s.cam.U * (2.f * p_x + -1.f) + s.cam.V * (-2.f * p_y + 1.f) + s.cam.W;
    const float dl = length(dir);
    dir = dl > 0.f ? dir / dl : V3f(0.f);
    // reparameterizeScreenSample takes the generator by value (WarpedAreaReparam.slang:305).
    if (P.useWAR) traceAux(s, tr, sg, s.cam.pos, dir, P.kappaS, P.aux, P.anti, r.screen);
    V3f o = s.cam.pos;
    Hit h = tr.closest(o, dir, 0.f, std::numeric_limits<float>::max());
    if (!h.valid()) return false;
    V3f thp(1.f), prevNormal(0.f);
    float prevPdf = 0.f;
    for (uint32_t len = 0;; len++) {
        VertexRec vr;
        vr.tri = h.tri;
        V3f posW, bary, faceN, normalW;
        computeIntersectionAD(s, h, o, dir, posW, bary, faceN, normalW);
        Hit th = h; th.u = bary.y; th.v = bary.z;
        ShadingData sd = loadShadingData(s, th, dir);
        sd.V = -dir;
        sd.frame = frameFromNormal(normalW);
        MatInst mi = makeMatInst(s, sd);
        vr.mat = sd.mat;
        vr.front = sd.frontFacing;
        vr.flip = !sd.frontFacing && matDoubleSided(s.mat(sd.mat));
        vr.faceN = widen(sd.faceN);
        const bool isPrimary = len == 0;
        if ((isPrimary || !P.nee || P.mis) && anyPositive(mi.emission)) {
            float misW = 1.f;
            if (!isPrimary && P.nee && P.mis) {
                const float lp = lightPdf(s, LS_UNIFORM, o, prevNormal, s.triToLight[h.tri], sd.posW, orientedFaceN(sd));
                misW = evalMIS(1.f, prevPdf, 1.f, lp);
            }
            vr.emit = true;
            vr.E = widen(mi.emission);
            vr.misE = misW;
        }
        if (len >= P.maxBounces + 1) {
            r.v.push_back(vr);
            break;
        }
        const V3f faceO = sd.frontFacing ? sd.faceN : -sd.faceN;
        if (P.nee) {
            LightSample ls;
            if (sampleLight(s, LS_UNIFORM, posW, normalW, sg, ls)) {
                V3f Li = ls.pdf > 0.f ? ls.Le / ls.pdf : V3f(0.f);
                V3f lightPos = computeRayOrigin(ls.posW, ls.normalW);
                V3f origin = computeRayOrigin(posW, normalW);
                V3f toLight = lightPos - origin;
                float distance = length(toLight);
                V3f ldir = normalize(toLight);
                if (anyPositive(Li) && tr.visible(origin, ldir, distance)) {
                    V3f wo = ls.posW - posW;
                    wo = wo / length(wo);
                    float misW = 1.f;
                    if (P.mis) {
                        float scatterPdf = miEvalPdf(sd, mi, ldir);
                        if (!P.bsdfSampling) {
                            V3f wiL = mi.sf.toLocal(sd.V), woL = mi.sf.toLocal(ldir);
                            scatterPdf = (mi.d.type == MAT_PBRT_DIFFUSE) ? scatterPdf
                                         : (validHemisphere(sd, wiL, woL, ldir) && std::min(wiL.z, woL.z) >= kMinCosTheta ? woL.z * Consts<float>::inv_pi : 0.f);
                        }
                        misW = evalMIS(1.f, ls.pdf, 1.f, scatterPdf);
                    }
                    vr.lit = true;
                    vr.lightPos = widen(ls.posW);
                    vr.lightN = widen(ls.normalW);
                    vr.Li = widen(Li);
                    vr.misN = misW;
                    // computeNEE (PTUtils.slang:211-249): the reparameterized ray starts at the
                    // view-side offset of the hit point, along wo.
                    if (P.useWAR) traceAux(s, tr, sg, computeRayOrigin(posW, faceO), wo, P.kappa, P.aux, P.anti, vr.nee);
                }
            }
        }
        BsdfSample bs;
        if (!(miSample(sd, mi, sg, bs, P.bsdfSampling) && bs.pdf > 0.f)) {
            r.v.push_back(vr);
            break;
        }
        const V3f rayO = computeRayOrigin(posW, faceO);
        const V3f f = miEval(sd, mi, bs.wo);
        vr.scat = true;
        vr.woS = widen(bs.wo);
        vr.oS0 = widen(rayO);
        vr.pdfS = bs.pdf;
        // generateScatterRay (PTUtils.slang:136-172): the warp follows the BSDF sample.
        if (P.useWAR) traceAux(s, tr, sg, rayO, bs.wo, P.kappa, P.aux, P.anti, vr.sc);
        r.v.push_back(vr);
        thp = thp * (f / bs.pdf);
        prevNormal = orientedFaceN(sd);
        prevPdf = bs.pdf;
        if (!anyPositive(thp)) break;
        h = tr.closest(rayO, bs.wo, 0.f, std::numeric_limits<float>::max());
        if (!h.valid()) break;
        o = rayO;
        dir = bs.wo;
    }
    return true;
}

// BSDF value of a vertex's material (StandardMaterialInstance.slang:63-74, PBRT diffuse) in double
// precision for a perturbed shading frame (normal N, view V) and direction wo.
V3d evalBsdf(const Scene& s, const VertexRec& r, V3d N, V3d V, V3d wo) {
    const float* m = s.mat(r.mat);
    Frame<double> f = frameFromNormal(N);
    if (r.flip) f.N = -f.N;
    const V3d wiL = f.toLocal(V), woL = f.toLocal(wo);
    const V3d faceN = r.faceN;
    const bool woTop = dot(wo, faceN) >= 0.0;
    BsdfData<double> d;
    d.type = int(m[M_TYPE]);
    V3d base(m[M_BASE], m[M_BASE + 1], m[M_BASE + 2]);
    if (d.type == MAT_PBRT_DIFFUSE) {
        if (std::min(wiL.z, woL.z) < double(kMinCosTheta) || r.front != woTop) return V3d(0.0);
        return base * std::max(woL.z, 0.0) * Consts<double>::inv_pi;
    }
    const bool isT = wiL.z * woL.z < 0.0;
    if (std::min(wiL.z, std::fabs(woL.z)) < double(kMinCosTheta)) return V3d(0.0);
    if (isT ? (r.front == woTop) : (r.front != woTop)) return V3d(0.0);
    const double ior = m[M_IOR], fr = (ior - 1.0) / (ior + 1.0), F0 = fr * fr, metal = m[M_METAL];
    d.eta = r.front ? 1.0 / ior : ior;
    d.diffTrans = m[M_DIFF_TRANS];
    d.specTrans = m[M_SPEC_TRANS];
    if (d.diffTrans > 0.0 || d.specTrans > 0.0) {
        const V3d t(m[M_TRANS], m[M_TRANS + 1], m[M_TRANS + 2]);
        d.transmission = m[M_THIN] != 0.f ? t : V3d(std::sqrt(t.x), std::sqrt(t.y), std::sqrt(t.z));
    }
    d.diffuse = base + (V3d(0.0) - base) * metal;
    d.specular = V3d(F0) + (base - V3d(F0)) * metal;
    d.roughness = m[M_ROUGH];
    d.metallic = metal;
    StdBsdf<double> b;
    b.init(wiL, d);
    return b.eval(wiL, woL);
}

// det(I + J(theta) - J(0)) of a ray warp and its direction V(theta) = u(theta) - u(0).
void rayWarp(const Perturbed& P, const Perturbed& P0, const Warp& w, V3d origin, V3d origin0, V3d dir, V3d dir0, V3d& V,
             double& detJ) {
    V = V3d(0.0);
    detJ = 1.0;
    if (!w.used || !w.valid) return;
    const V3d e[3] = {V3d(1.0, 0.0, 0.0), V3d(0.0, 1.0, 0.0), V3d(0.0, 0.0, 1.0)};
    const WarpEval a = evalWarp(P, w, origin, dir, e, 3);
    const WarpEval b = evalWarp(P0, w, origin0, dir0, e, 3);
    V = a.u - b.u;
    V3d c[3];
    for (int k = 0; k < 3; k++) c[k] = e[k] + a.du[k] - b.du[k];
    detJ = det3(c);
}

// F(theta): dLdI . radiance of the recorded path (WARDiffPathTracer.rt.slang:75-111,
// PTUtils.slang:136-341 with the decisions held). Each vertex is attached to its incoming ray
// (AttachToRay: the ray's origin follows the previous vertex, computeRayOrigin's identity
// derivative, and its direction is the held sample plus the previous vertex's warp), so moving a
// vertex moves every later vertex of the path.
double pathValue(const Scene& s, const Params& Pp, const PathRec& r, const Perturbed& P, const Perturbed& P0) {
    const CamD cam{widen(s.cam.pos), widen(s.cam.U), widen(s.cam.V), widen(s.cam.W)};
    // Screen warp (reparameterizeScreenSample): posS + V_s and det(I + J_s(theta) - J_s(0)).
    double vx = 0.0, vy = 0.0, detS = 1.0;
    if (r.screen.used && r.screen.valid) {
        const V3d r0 = camRaw(cam, r.psx, r.psy, Pp.W, Pp.H);
        const double rl = length(r0);
        const V3d w = rl > 0.0 ? r0 / rl : V3d(0.0);
        // tangents of the normalized camera direction along the two screen axes
        auto nw = [&](V3d v) { return rl > 0.0 ? (v - w * dot(w, v)) / rl : V3d(0.0); };
        const V3d t[2] = {nw(cam.U * (2.0 / double(Pp.W))), nw(cam.V * (-2.0 / double(Pp.H)))};
        double sxA, syA, sxB, syB, jA[2][2], jB[2][2];
        const WarpEval a = evalWarp(P, r.screen, cam.pos, w, t, 2);
        const WarpEval b = evalWarp(P0, r.screen, cam.pos, w, t, 2);
        for (int k = 0; k < 2; k++) {
            double dx, dy;
            camScreen(cam, a.u, a.du[k], Pp.W, Pp.H, sxA, syA, dx, dy);
            jA[0][k] = dx; jA[1][k] = dy;
            camScreen(cam, b.u, b.du[k], Pp.W, Pp.H, sxB, syB, dx, dy);
            jB[0][k] = dx; jB[1][k] = dy;
        }
        vx = sxA - sxB;
        vy = syA - syB;
        const double m00 = 1.0 + jA[0][0] - jB[0][0], m01 = jA[0][1] - jB[0][1];
        const double m10 = jA[1][0] - jB[1][0], m11 = 1.0 + jA[1][1] - jB[1][1];
        detS = m00 * m11 - m01 * m10;
    }
    const double sx = r.psx + vx, sy = r.psy + vy;
    // evalGaussian(pixel - center) / detach(...) (sigma 0.5)
    const double alpha = -1.0 / (2.0 * 0.25);
    const double g = std::exp(alpha * ((sx - r.pcx) * (sx - r.pcx) + (sy - r.pcy) * (sy - r.pcy)));
    const double g0 = std::exp(alpha * ((r.psx - r.pcx) * (r.psx - r.pcx) + (r.psy - r.pcy) * (r.psy - r.pcy)));
    // AttachToRay (loadVertexPosW with a differentiable ray): hit point and interpolated normal.
    auto attach = [&](const Perturbed& Q, uint32_t tri, V3d o, V3d d, V3d& x, V3d& n) {
        V3d v[3];
        Q.tri(tri, v);
        const double t = computeRayDistanceToTriangle(v, o, d);
        const V3d p = o + d * t;
        const V3d b = computeBarycentrics(v, p);
        x = v[0] * b.x + v[1] * b.y + v[2] * b.z;
        const uint32_t* id = &s.idx[3 * tri];
        n = Q.nrm(id[0]) * b.x + Q.nrm(id[1]) * b.y + Q.nrm(id[2]) * b.z;
    };
    V3d thp(detS * g / g0);
    V3d L(0.0);
    V3d o = cam.pos, o0 = cam.pos;
    V3d d = normalizeSafeD(camRaw(cam, sx, sy, Pp.W, Pp.H)), d0 = normalizeSafeD(camRaw(cam, r.psx, r.psy, Pp.W, Pp.H));
    for (const VertexRec& v : r.v) {
        V3d x, n, x0, n0;
        attach(P, v.tri, o, d, x, n);
        attach(P0, v.tri, o0, d0, x0, n0);
        const V3d view = -d;
        if (v.emit) L = L + thp * v.E * v.misE;
        // The rays leaving the vertex start at its offset point, which follows the vertex
        // (fwd/bwd_computeRayOrigin: identity derivative).
        if (v.lit) {
            // computeNEE: wo = normalize(y - x) toward the held light point y, and the area factor
            // a2s / detach(a2s). With neeHeld the direction stays at its sampled value and the
            // factor is 1.
            V3d wo0 = v.lightPos - x0;
            const double dist0 = length(wo0);
            wo0 = wo0 / dist0;
            V3d wo = Pp.neeHeld ? wo0 : v.lightPos - x;
            const double dist = Pp.neeHeld ? dist0 : length(wo);
            if (!Pp.neeHeld) wo = wo / dist;
            const V3d oN = v.nee.origin0 + (x - x0);
            V3d V;
            double detJ;
            rayWarp(P, P0, v.nee, oN, v.nee.origin0, wo, wo0, V, detJ);
            const V3d rd = wo + V;
            const V3d f = evalBsdf(s, v, n, view, rd);
            // A light sample in the receiver's tangent plane (a hit point flush with the light) has a
            // zero BSDF value and an unbounded area factor; its contribution is zero (in FP32 the
            // factor cancels to 1 at the base point).
            const double a2s = dist * dist / dot(v.lightN, -rd), a2s0 = dist0 * dist0 / dot(v.lightN, -wo0);
            if (anyPositive(f) && a2s0 > 0.0 && std::isfinite(a2s0))
                L = L + thp * v.Li * f * (Pp.neeHeld ? 1.0 : a2s / a2s0) * (v.misN * detJ);
        }
        if (!v.scat) break;
        const V3d oS = v.oS0 + (x - x0);
        V3d V;
        double detJ;
        rayWarp(P, P0, v.sc, oS, v.oS0, v.woS, v.woS, V, detJ);
        const V3d rd = v.woS + V;
        const V3d f = evalBsdf(s, v, n, view, rd);
        thp = thp * f * (detJ / v.pdfS);
        o = oS;
        o0 = v.oS0;
        d = rd;
        d0 = v.woS;
    }
    return dot(r.dL, L);
}

}  // namespace
}  // namespace fo

using namespace fo;

extern "C" {
// Directional derivatives of the WARDiffPathTracer loss sum_pixels dLdI . I at maxBounces
// (useWAR = 1) along nprobe perturbation fields, and the loss itself. probes: [nprobe][nV][6]
// (position then normal perturbation of each scene vertex). The image I is the per-pixel sum
// over the spp samples with seeds frame0 + k (the adjoint's per-sample accumulation, not the
// average). flags: bit0 useBSDFSampling, bit1 useNEE, bit2 useMIS, bit3 useAntithetic, bit5 useWAR,
// bit6 the held NEE direction (the product's useSolidAngleNEEWarp extension).
// out[k] = sum_paths dLdI . dF/dtheta along probe k by central differences with step h (checked
// against step h/2 per path, see below); value
// receives sum_paths dLdI . F(0) (the primal, for the record check).
void fo_war_geom_probe_b(void* hs, uint32_t W, uint32_t H, uint32_t spp, uint32_t frame0, uint32_t maxBounces, uint32_t aux,
                         float log10Kappa, float log10KappaScreen, uint32_t flags, const float* dLdI, uint32_t nprobe,
                         const float* probes, double h, int threads, double* out, double* value) {
    const Scene& s = *static_cast<Scene*>(hs);
    Params P{W, H, aux, float(std::pow(10.0, double(log10Kappa))), float(std::pow(10.0, double(log10KappaScreen))),
             (flags & 32) != 0, (flags & 1) != 0, (flags & 2) != 0, (flags & 4) != 0, (flags & 8) != 0,
             (flags & 64) != 0, maxBounces};
    const size_t nV = s.pos.size();
    // Position and normal perturbation of each probe, [nV][3] each.
    std::vector<std::vector<float>> dps(nprobe, std::vector<float>(3 * nV)), dns(nprobe, std::vector<float>(3 * nV));
    for (uint32_t q = 0; q < nprobe; q++) {
        const float* pr = probes + size_t(q) * nV * 6;
        for (size_t v = 0; v < nV; v++)
            for (int c = 0; c < 3; c++) {
                dps[q][3 * v + c] = pr[6 * v + c];
                dns[q][3 * v + c] = pr[6 * v + 3 + c];
            }
    }
    unsigned nth = threads > 0 ? unsigned(threads) : std::max(1u, std::thread::hardware_concurrency());
    // Per-row sums, added in row order afterwards: the result does not depend on the scheduling.
    std::vector<std::vector<double>> acc(H, std::vector<double>(nprobe + 1, 0.0));
    std::atomic<uint32_t> next{0};
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nth; t++)
        pool.emplace_back([&, t] {
            (void)t;
            Tracer tr{s, false, nullptr, nullptr};
            const Perturbed P0{s, nullptr, nullptr, 0.0};
            for (uint32_t row; (row = next.fetch_add(1)) < H;) {
                std::vector<double>& a = acc[row];
                for (uint32_t x = 0; x < W; x++) {
                    const float* dl = dLdI + 3 * (size_t(row) * W + x);
                    if (dl[0] == 0.f && dl[1] == 0.f && dl[2] == 0.f) continue;
                    for (uint32_t k = 0; k < spp; k++) {
                        PathRec r;
                        r.dL = V3d(dl[0], dl[1], dl[2]);
                        if (!recordPath(s, P, x, row, frame0 + k, tr, r)) continue;
                        const double f0 = pathValue(s, P, r, P0, P0);
                        a[nprobe] += f0;
                        for (uint32_t q = 0; q < nprobe; q++) {
                            auto cd = [&](double step) {
                                const Perturbed Pp{s, dps[q].data(), dns[q].data(), step}, Pm{s, dps[q].data(), dns[q].data(), -step};
                                return (pathValue(s, P, r, Pp, P0) - pathValue(s, P, r, Pm, P0)) / (2.0 * step);
                            };
                            // F is smooth except at the evaluation's own branch points (cosine
                            // cut-offs, hemisphere tests) where it jumps; the held decisions are those
                            // taken at theta = 0, so a difference across a jump is discarded: if the
                            // steps h and h/2 disagree, steps 1000 times smaller are used, and a path
                            // that still jumps contributes the derivative of its branch at 0 (zero if
                            // F(0) = 0, the smaller step otherwise).
                            double d1 = cd(h), d2 = cd(0.5 * h);
                            if (std::fabs(d1 - d2) > 1e-4 * std::fabs(d2) + 1e-9 * (std::fabs(f0) + 1.0) / h) {
                                d1 = cd(1e-3 * h);
                                d2 = cd(5e-4 * h);
                                if (std::fabs(d1 - d2) > 1e-3 * std::fabs(d2) + 1e-6 * (std::fabs(f0) + 1.0) / h)
                                    d2 = f0 == 0.0 ? 0.0 : d2;
                            }
                            a[q] += d2;
                        }
                    }
                }
            }
        });
    for (auto& th : pool) th.join();
    for (uint32_t q = 0; q <= nprobe; q++) {
        double sum = 0.0;
        for (uint32_t y = 0; y < H; y++) sum += acc[y][q];
        if (q < nprobe) out[q] = sum;
        else if (value) *value = sum;
    }
}
// The same at maxBounces = 0 (direct illumination).
void fo_war_geom_probe(void* hs, uint32_t W, uint32_t H, uint32_t spp, uint32_t frame0, uint32_t aux, float log10Kappa,
                       float log10KappaScreen, uint32_t flags, const float* dLdI, uint32_t nprobe, const float* probes,
                       double h, int threads, double* out, double* value) {
    fo_war_geom_probe_b(hs, W, H, spp, frame0, 0, aux, log10Kappa, log10KappaScreen, flags, dLdI, nprobe, probes, h, threads,
                        out, value);
}
}  // extern "C"
