// shading.hpp -- hit geometry, BSDFs, lights and camera on the device.
//
// Each function restates one reference routine (cited) with the same
// operation order so that results match the CPU renderer bit for bit.
#pragma once
#include "device.hpp"

namespace bnd {

struct Geom { V3 p, n, t, b; };   // world position and the transformed ONB

// OrthonormalBasis(n) (Base/Primitive.fs:15-23)
__device__ __forceinline__ void onb_from_n(V3 n, V3& t, V3& b) {
    V3 axis = fabsf(n.x) > 0.1f ? mk(0, 1, 0) : mk(1, 0, 0);
    t = normalize(cross(n, axis));
    b = cross(n, t);
}

// Geometry of a closest hit (inst, prim, t) along world ray r: the object-space
// intersection point and normal of Mesh.fs:76-78 / Sphere.fs:52-61 mapped by
// LocalGeometry.Transform (Base/Primitive.fs:34-38, 57-58).
__device__ __forceinline__ Geom hit_geometry(const DevScene& S, const Ray& r, int inst, int prim, float t) {
    const DevInst& I = S.insts[inst];
    const Xf w2o = I.w2o;
    Ray o = {xf_point(w2o, r.o), xf_dir(w2o, r.d)};
    V3 n;
    const int4 info = I.info;
    if (info.x < 0) {
        V3 p = point_at(o, t);
        n = normalize(p);
        if (prim == 0 && dot(n, o.d) > 0.0f) n = -n;
    } else {
        Tri tr = load_tri(S.tris, info.y + prim);
        n = normalize(cross(tr.p1 - tr.p0, tr.p2 - tr.p0));
    }
    V3 p = point_at(o, t);
    V3 lt, lb;
    onb_from_n(n, lt, lb);
    const Xf o2w = I.o2w;
    Geom g;
    g.p = xf_point(o2w, p);
    g.t = normalize(xf_dir(o2w, lt));
    g.b = normalize(xf_dir(o2w, lb));
    g.n = cross(g.t, g.b);
    return g;
}
__device__ __forceinline__ V3 to_local(const Geom& g, V3 v) { return {dot(v, g.t), dot(v, g.b), dot(v, g.n)}; }
__device__ __forceinline__ V3 to_world(const Geom& g, V3 v) { return v.x * g.t + v.y * g.b + v.z * g.n; }

// ------------------------------------------------------------- materials
struct Mat { int kind; V3 albedo; float ior, alpha, metallic, roughness; };
__device__ __forceinline__ Mat load_mat(const DevScene& S, int m) {
    float4 a = S.mats[2 * m], b = S.mats[2 * m + 1];
    return {__float_as_int(a.x), {a.y, a.z, a.w}, b.x, b.y, b.z, b.w};
}
struct BSDF { V3 f; float pdf; };

__device__ __forceinline__ float pbr_lambda(float alpha, V3 w) {         // PBR.fs:13-19
    float s2 = fmaf(w.x, w.x, w.y * w.y);
    if (s2 == 0.0f) return 0.0f;
    float t2 = s2 / (w.z * w.z);
    float a2t2 = alpha * alpha * t2;
    return (-1.0f + sqrtf(1.0f + a2t2)) / 2.0f;
}
__device__ __forceinline__ BSDF pbr_eval(const Mat& m, V3 wo, V3 wi) {   // PBR.fs:37-49
    V3 wh = normalize(wo + wi);
    float c = fabsf(wh.z);
    float x = 1.0f + fmaf(m.alpha, m.alpha, -1.0f) * c * c;
    float d = m.alpha * m.alpha / (kPi * (x * x));
    float g = 1.0f / (1.0f + pbr_lambda(m.alpha, wo) + pbr_lambda(m.alpha, wi));
    float spec = d * g / (4.0f * fabsf(wo.z));
    float cw = dot(wo, wh);
    float k = 1.0f - cw;
    float k2 = k * k;
    V3 fc = m.albedo + (mk(1, 1, 1) - m.albedo) * (k2 * k2 * k);
    V3 metal = spec * fc;
    const float f0 = 0.04f;
    float f = f0 + (1.0f - f0) * k2 * k2 * k;
    V3 diff = m.albedo * fabsf(wi.z) / kPi;
    float s1 = 1.0f - f;
    V3 sf = mk(spec, spec, spec) * f;
    V3 inner = {fmaf(diff.x, s1, sf.x), fmaf(diff.y, s1, sf.y), fmaf(diff.z, s1, sf.z)};
    float sm = 1.0f - m.metallic;
    V3 mm = metal * m.metallic;
    V3 bsdf = {fmaf(inner.x, sm, mm.x), fmaf(inner.y, sm, mm.y), fmaf(inner.z, sm, mm.z)};
    float pa = d * fabsf(wh.z) / (4.0f * cw);
    float pb = fabsf(wi.z) / kPi;
    float tt = 0.5f * (1.0f - m.metallic);
    return {bsdf, fmaf(pa, 1.0f - tt, pb * tt)};
}
__device__ __forceinline__ BSDF mat_eval(const Mat& m, V3 wo, V3 wi) {
    if (m.kind == 0) {                                                    // Lambertian.fs:11-16
        if (wi.z * wo.z < 0.0f || min_n(fabsf(wi.z), fabsf(wo.z)) < 1e-6f) return {{0, 0, 0}, 0.0f};
        float pdf = fabsf(wi.z) / kPi;
        return {m.albedo * pdf, pdf};
    }
    if (m.kind == 3) return pbr_eval(m, wo, wi);
    return {{0, 0, 0}, 0.0f};                                             // Mirror.fs:11, Dielectric.fs:13
}
__device__ __forceinline__ BSDF mat_sample(const Mat& m, V3 wo, float ul, V2 u, V3& wi) {
    if (m.kind == 0) {                                                    // Lambertian.fs:18-26
        float ct = sqrtf(u.x), st = sqrtf(1.0f - u.x);
        float ph = 2.0f * kPi * u.y;
        float sp = sin_f(ph), cp = cos_f(ph);
        V3 w = {st * cp, st * sp, ct};
        float pdf = w.z / kPi;
        wi = wo.z > 0.0f ? w : -w;
        return {m.albedo * pdf, pdf};
    }
    if (m.kind == 1) {                                                    // Mirror.fs:13-14
        wi = {-wo.x, -wo.y, wo.z};
        return {m.albedo, 1.0f};
    }
    if (m.kind == 2) {                                                    // Dielectric.fs:15-31
        float ior = wo.z > 0.0f ? 1.0f / m.ior : m.ior;
        float c2t = 1.0f - ior * ior * fmaf(-wo.z, wo.z, 1.0f);
        if (c2t <= 0.0f) { wi = {-wo.x, -wo.y, wo.z}; return {m.albedo, 1.0f}; }
        float a = ior - 1.0f, b = ior + 1.0f;
        float r0 = a * a / (b * b);
        float c = 1.0f - fabsf(wo.z);
        float c2 = c * c;
        float r = r0 + (1.0f - r0) * c2 * c2 * c;
        if (ul < r) { wi = {-wo.x, -wo.y, wo.z}; return {r * m.albedo, r}; }
        float ct = sqrtf(c2t);
        wi = {-wo.x * ior, -wo.y * ior, -copysignf(ct, wo.z)};
        return {(1.0f - r) * m.albedo, 1.0f - r};
    }
    // PBR.fs:50-64
    if (ul < 1.0f - 0.5f * (1.0f - m.metallic)) {
        float th = atan_f(m.alpha * sqrtf(u.x / (1.0f - u.x)));
        float ph = 2.0f * kPi * u.y;
        float st = sin_f(th), ct = cos_f(th), sp = sin_f(ph), cp = cos_f(ph);
        V3 wh = {st * cp, st * sp, ct};
        wi = 2.0f * dot(wo, wh) * wh - wo;
    } else {
        float ct = sqrtf(u.x), st = sqrtf(1.0f - u.x);
        float ph = 2.0f * kPi * u.y;
        float sp = sin_f(ph), cp = cos_f(ph);
        wi = {st * cp, st * sp, ct};
    }
    return pbr_eval(m, wo, wi);
}

// ------------------------------------------------------------------ lights
// Light record per light (TLAS order): (instance, first triangle, triangle
// count, light definition).
__device__ __forceinline__ V3 light_emit(float4 def, float woz) {        // Light.fs:49-52
    if (fabsf(woz) > 1e-6f && (woz > 0.0f || def.w != 0.0f)) return xyz(def);
    return {0, 0, 0};
}
struct LightSample { V3 p, L; float pdf; V3 wi; int ldef; float facing; };

// ------------------------------------------------------------ derivatives
// d f / d albedo (per channel; f_c depends on albedo_c only) and d f / d alpha
// (per channel) of a BSDF value at fixed directions, for the path-replay
// adjoint.  The value and pdf are those of mat_eval / mat_sample, bit for bit.
struct BSDFGrad { V3 f; float pdf; V3 dalb; V3 dalpha; float dpdf; };   // dpdf: d pdf / d alpha

__device__ __forceinline__ float pbr_dlambda(float alpha, V3 w) {
    float s2 = fmaf(w.x, w.x, w.y * w.y);
    if (s2 == 0.0f) return 0.0f;
    float t2 = s2 / (w.z * w.z);
    return 0.5f * alpha * t2 / sqrtf(1.0f + alpha * alpha * t2);
}
__device__ __forceinline__ BSDFGrad pbr_eval_grad(const Mat& m, V3 wo, V3 wi) {
    const BSDF e = pbr_eval(m, wo, wi);
    V3 wh = normalize(wo + wi);
    float c = fabsf(wh.z);
    float a = m.alpha;
    float x = 1.0f + (a * a - 1.0f) * c * c;
    float d = a * a / (kPi * (x * x));
    float dd = 2.0f * a / (kPi * x * x) * (1.0f - 2.0f * a * a * c * c / x);
    float g = 1.0f / (1.0f + pbr_lambda(a, wo) + pbr_lambda(a, wi));
    float dg = -g * g * (pbr_dlambda(a, wo) + pbr_dlambda(a, wi));
    float spec = d * g / (4.0f * fabsf(wo.z));
    float dspec = (dd * g + d * dg) / (4.0f * fabsf(wo.z));
    float k = 1.0f - dot(wo, wh);
    float k5 = k * k * k * k * k;
    float F = 0.04f + 0.96f * k5;
    float sm = 1.0f - m.metallic;
    float base = sm * (1.0f - F) * fabsf(wi.z) / kPi + m.metallic * spec * (1.0f - k5);
    V3 fc = m.albedo + (mk(1, 1, 1) - m.albedo) * k5;
    BSDFGrad r;
    r.f = e.f;
    r.pdf = e.pdf;
    r.dalb = mk(base, base, base);
    r.dalpha = dspec * (mk(sm * F, sm * F, sm * F) + m.metallic * fc);
    r.dpdf = dd * c / (4.0f * dot(wo, wh)) * (1.0f - 0.5f * (1.0f - m.metallic));
    return r;
}
__device__ __forceinline__ BSDFGrad mat_eval_grad(const Mat& m, V3 wo, V3 wi) {
    if (m.kind == 3) return pbr_eval_grad(m, wo, wi);
    const BSDF e = mat_eval(m, wo, wi);
    BSDFGrad r;
    r.f = e.f;
    r.pdf = e.pdf;
    r.dalb = m.kind == 0 && e.pdf != 0.0f ? mk(e.pdf, e.pdf, e.pdf) : mk(0, 0, 0);
    r.dalpha = mk(0, 0, 0);
    r.dpdf = 0.0f;
    return r;
}
__device__ __forceinline__ BSDFGrad mat_sample_grad(const Mat& m, V3 wo, float ul, V2 u, V3& wi) {
    const BSDF e = mat_sample(m, wo, ul, u, wi);
    if (m.kind == 3) {
        BSDFGrad r = pbr_eval_grad(m, wo, wi);
        r.f = e.f;
        r.pdf = e.pdf;
        return r;
    }
    BSDFGrad r;
    r.f = e.f;
    r.pdf = e.pdf;
    // Lambertian: albedo * pdf; mirror / total reflection: albedo; dielectric:
    // r * albedo or (1 - r) * albedo with the returned pdf r or 1 - r.
    const float k = m.kind == 0 ? e.pdf : (m.kind == 1 ? 1.0f : e.pdf);
    r.dalb = mk(k, k, k);
    r.dalpha = mk(0, 0, 0);
    r.dpdf = 0.0f;
    return r;
}
// d alpha / d roughness of Alpha = Max(roughness^2, 1e-3) (PBR.fs:11)
__device__ __forceinline__ float dalpha_drough(const Mat& m) {
    float r2 = m.roughness * m.roughness;
    return r2 > 1e-3f ? 2.0f * m.roughness : 0.0f;
}

// UniformLightSampler.Sample (Uniform.fs:13-29) over MeshInstance.Sample
// (Mesh.fs:289-298, AliasTable.fs:52-62, Triangle.Sample Mesh.fs:89-99) or
// SphereInstance.Sample (Sphere.fs:90-113).
__device__ __forceinline__ LightSample light_sample(const DevScene& S, V3 p, float uSel, V2 u) {
    const int N = S.n_lights;
    float us = uSel * (float)N;
    int id = sat_int(us);
    id = id < N - 1 ? id : N - 1;
    us = us - (float)id;
    const int4 rec = S.lrec[id];
    const DevInst& I = S.insts[rec.x];
    const Xf o2w = I.o2w;
    V3 lp, ln;
    float pdfS;
    if (I.info.x < 0) {
        const float radius = I.lo.w;
        float a = 2.0f * kPi * u.x;
        float st = sin_f(a), ct = cos_f(a);
        float cp = fmaf(-2.0f, u.y, 1.0f);
        float sp = sqrtf(fmaf(-cp, cp, 1.0f));
        V3 nn = {ct * sp, st * sp, cp};
        V3 ot, ob;
        onb_from_n(nn, ot, ob);
        lp = xf_point(o2w, nn * radius);
        V3 nw = cross(xf_dir(o2w, ot), xf_dir(o2w, ob));
        float invJ = 1.0f / length(nw);
        ln = invJ * nw;
        pdfS = invJ / (4.0f * kPi * radius * radius);
    } else {
        const int n = rec.z;
        float uu = us * (float)n;
        int i = sat_int(uu);
        i = i < n - 1 ? i : n - 1;
        i = i < 0 ? 0 : i;
        uu = uu - (float)i;
        int idx = uu < S.alias_prob[rec.y + i] ? i : S.alias_idx[rec.y + i];
        float pdfTri = S.alias_pdf[rec.y + idx];
        Tri tr = load_tri(S.tris, rec.y + idx);
        V3 p0 = xf_point(o2w, tr.p0), p1 = xf_point(o2w, tr.p1), p2 = xf_point(o2w, tr.p2);
        V2 uv = u.x < u.y ? V2{0.5f * u.x, fmaf(-0.5f, u.x, u.y)} : V2{fmaf(-0.5f, u.y, u.x), 0.5f * u.y};
        lp = uv.x * p1 + uv.y * p2 + (1.0f - uv.x - uv.y) * p0;
        V3 nc = cross(p1 - p0, p2 - p0);
        float pdfArea = 2.0f / length(nc);
        ln = (0.5f * pdfArea) * nc;
        pdfS = pdfTri * pdfArea;
    }
    V3 wo = normalize(p - lp);
    float cosWo = dot(ln, wo);
    V3 dv = p - lp;
    float dist2 = dot(dv, dv);
    LightSample s;
    s.p = lp;
    const float4 def = S.lightdefs[rec.w];
    const float woz = dot(wo, ln);
    s.L = light_emit(def, woz);
    s.pdf = dist2 * pdfS / (max_n(fabsf(cosWo), 1e-6f) * (float)N);
    s.wi = -wo;
    s.ldef = rec.w;
    s.facing = fabsf(woz) > 1e-6f && (woz > 0.0f || def.w != 0.0f) ? 1.0f : 0.0f;
    return s;
}

// UniformLightSampler.Eval (Uniform.fs:40-49) at a BSDF-sampled hit; the
// mesh EvalPDF (Mesh.fs:300-304, triangle 0 because the tag is lost) is a
// per-instance constant computed on the host.
__device__ __forceinline__ void light_eval(const DevScene& S, int inst, V3 o, const Geom& g, V3& L, float& pdf) {
    const DevInst& I = S.insts[inst];
    V3 wo = normalize(o - g.p);
    float cosWo = dot(g.n, wo);
    float pdfS;
    if (I.info.x < 0) {                                                  // Sphere.fs:115-126
        const Xf w2o = I.w2o;
        float invJ = length(cross(xf_dir(w2o, g.t), xf_dir(w2o, g.b)));
        float r = I.lo.w;
        pdfS = invJ / (4.0f * kPi * r * r);
    } else {
        pdfS = I.hi.w;
    }
    V3 dv = o - g.p;
    float dist2 = dot(dv, dv);
    L = light_emit(S.lightdefs[I.info.w], dot(wo, g.n));
    pdf = dist2 * pdfS / (max_n(fabsf(cosWo), 1e-6f) * (float)S.n_lights);
}

// ------------------------------------------------------------------ camera
// CameraBase.GeneratePrimaryRay (Base/Camera.fs:12-20) over Pinhole.fs:12-27
// and ThinLens.fs:12-35.  The viewport corner and pixel deltas are computed
// once on the host with the same operations.
__device__ __forceinline__ Ray camera_ray(const DevScene& S, int x, int y, V2 uPix, V2 uLens) {
    V3 pl = xyz(S.ul) + ((float)x + uPix.x) * xyz(S.du) + ((float)y + uPix.y) * xyz(S.dv);
    Ray r = {{0, 0, 0}, normalize(pl)};
    const float aperture = S.du.w;
    if (S.cam_kind == 1 && aperture > 0.0f) {
        V2 u = {uLens.x * 2.0f - 1.0f, uLens.y * 2.0f - 1.0f};
        V2 d = {0.0f, 0.0f};
        if (u.x == 0.0f || u.y == 0.0f) {
            d = {0.0f, 0.0f};
        } else if (fabsf(u.x) > fabsf(u.y)) {
            float rr = u.x, th = kPi / 4.0f * (u.y / u.x);
            d = {rr * cos_f(th), rr * sin_f(th)};
        } else {
            float rr = u.y, th = kPi / 2.0f - kPi / 4.0f * (u.x / u.y);
            d = {rr * cos_f(th), rr * sin_f(th)};
        }
        V3 o = r.o + mk(aperture * d.x, aperture * d.y, 0.0f);
        V3 dir = normalize(point_at(r, S.dv.w) - o);
        r = {o, dir};
    }
    Ray w;
    w.o = xf_point(S.c2w, r.o);
    w.d = normalize(xf_dir(S.c2w, r.d));
    w.o = point_at(w, S.ul.w);
    return w;
}

}  // namespace bnd
