// Exact adjoint of the path tracer (device code; kernel and launch in adjoint.hip): gradients of every forward output with respect to
// every scene parameter (positions, rotations, scales, densities, SH coefficients,
// shading normals, materials, environment texels), with the path's discrete decisions
// (candidate sets, acceptance, escape/shadow/roulette draws) held fixed.
//
// No counterpart in the reference: its backward (referenceBwdOptix.cu:27-74,
// pathTracer.cuh:313-634) propagates material and environment gradients only, through
// forward-difference and disabled terms.  The per-hit geometry terms follow the same
// quantities as processHit (gaussianParticles.cuh:343-440; its hand-written adjoint for the
// 3DGRT tracer is processHitBwd, gaussianParticles.cuh:502-837), derived here from the
// forward expressions.
//
// One lane owns one path, as in the forward.  The kernel
//   1. replays the path with the forward's code and random draws and records, per path
//      vertex, the segment ray, its totals (transmittance, depth, shading normal, material,
//      radiance) and the uniforms of the BRDF sample taken there (AdjIO::scratch);
//   2. sweeps the vertices in reverse, carrying the adjoints of the throughput and of the
//      segment ray: contribution terms (escape to the environment, INDIRECT radiance,
//      NEE with its MIS weight) -> the vertex's normalised quantities -> the segment's
//      hits, re-traced front to back with the segment totals giving every suffix sum ->
//      the BRDF sample that produced the segment -> the previous vertex.
// The local Jacobians of the per-vertex blocks (BRDF sampling and evaluation, environment
// lookup, SH in the ray direction) come from forward-mode jets run over the same
// expressions as the forward (templated copies below); the per-hit terms, which run once per
// accepted hit, are written out by hand.
#pragma once

#include "trace_path.hpp"

namespace ptir {
namespace adj {

// ---------------------------------------------------------------- forward-mode jets
template <int N>
struct Jet {
    float v;
    float d[N];
    __device__ __forceinline__ Jet() {}
    __device__ __forceinline__ Jet(float x) : v(x) {
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] = 0.f;
    }
};

#define JET_BIN(OP, VEXPR, DEXPR)                                                                         \
    template <int N>                                                                                      \
    __device__ __forceinline__ Jet<N> operator OP(const Jet<N>& a, const Jet<N>& b) {                     \
        Jet<N> r;                                                                                         \
        r.v = VEXPR;                                                                                      \
        _Pragma("unroll") for (int i = 0; i < N; ++i) r.d[i] = DEXPR;                                     \
        return r;                                                                                         \
    }
JET_BIN(+, a.v + b.v, a.d[i] + b.d[i])
JET_BIN(-, a.v - b.v, a.d[i] - b.d[i])
JET_BIN(*, a.v * b.v, a.d[i] * b.v + a.v * b.d[i])
#undef JET_BIN
template <int N>
__device__ __forceinline__ Jet<N> operator/(const Jet<N>& a, const Jet<N>& b) {
    Jet<N> r;
    r.v = a.v / b.v;
    const float ib = 1.f / b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
    return r;
}
template <int N> __device__ __forceinline__ Jet<N> operator-(const Jet<N>& a) {
    Jet<N> r;
    r.v = -a.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
    return r;
}
template <int N> __device__ __forceinline__ Jet<N> operator+(const Jet<N>& a, float b) { return a + Jet<N>(b); }
template <int N> __device__ __forceinline__ Jet<N> operator-(const Jet<N>& a, float b) { return a - Jet<N>(b); }
template <int N> __device__ __forceinline__ Jet<N> operator+(float a, const Jet<N>& b) { return Jet<N>(a) + b; }
template <int N> __device__ __forceinline__ Jet<N> operator-(float a, const Jet<N>& b) { return Jet<N>(a) - b; }
template <int N> __device__ __forceinline__ Jet<N> operator*(const Jet<N>& a, float b) {
    Jet<N> r;
    r.v = a.v * b;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b;
    return r;
}
template <int N> __device__ __forceinline__ Jet<N> operator*(float a, const Jet<N>& b) { return b * a; }
template <int N> __device__ __forceinline__ Jet<N> operator/(const Jet<N>& a, float b) { return a / Jet<N>(b); }
template <int N> __device__ __forceinline__ Jet<N> operator/(float a, const Jet<N>& b) { return Jet<N>(a) / b; }

__device__ __forceinline__ float val(float x) { return x; }
template <int N> __device__ __forceinline__ float val(const Jet<N>& x) { return x.v; }

// the derivative of a branch is the derivative of the branch its value takes (ties: first argument)
template <class S> __device__ __forceinline__ S smax(const S& a, const S& b) { return val(a) >= val(b) ? a : b; }
template <class S> __device__ __forceinline__ S smin(const S& a, const S& b) { return val(a) <= val(b) ? a : b; }

__device__ __forceinline__ float ssqrt(float x) { return sqrtf(x); }
template <int N> __device__ __forceinline__ Jet<N> ssqrt(const Jet<N>& x) {
    Jet<N> r;
    r.v = sqrtf(x.v);
    const float k = r.v > 0.f ? 0.5f / r.v : 0.f;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = x.d[i] * k;
    return r;
}
__device__ __forceinline__ float srsqrt(float x) { return rsqrtf(x); }
template <int N> __device__ __forceinline__ Jet<N> srsqrt(const Jet<N>& x) {
    Jet<N> r;
    r.v = rsqrtf(x.v);
    const float k = -0.5f * r.v * r.v * r.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = x.d[i] * k;
    return r;
}
__device__ __forceinline__ void ssincos(float x, float& s, float& c) { sincosf(x, &s, &c); }
template <int N> __device__ __forceinline__ void ssincos(const Jet<N>& x, Jet<N>& s, Jet<N>& c) {
    float sv, cv;
    sincosf(x.v, &sv, &cv);
    s.v = sv;
    c.v = cv;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        s.d[i] = cv * x.d[i];
        c.d[i] = -sv * x.d[i];
    }
}
__device__ __forceinline__ float satan2(float y, float x) { return atan2f(y, x); }
template <int N> __device__ __forceinline__ Jet<N> satan2(const Jet<N>& y, const Jet<N>& x) {
    Jet<N> r;
    r.v = atan2f(y.v, x.v);
    const float den = x.v * x.v + y.v * y.v;
    const float k = den > 0.f ? 1.f / den : 0.f;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = (x.v * y.d[i] - y.v * x.d[i]) * k;
    return r;
}
__device__ __forceinline__ float sasin(float x) { return asinf(x); }
template <int N> __device__ __forceinline__ Jet<N> sasin(const Jet<N>& x) {
    Jet<N> r;
    r.v = asinf(x.v);
    const float k = rsqrtf(fmaxf(1.f - x.v * x.v, 1e-30f));
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = x.d[i] * k;
    return r;
}
__device__ __forceinline__ float sfloor(float x) { return floorf(x); }
template <int N> __device__ __forceinline__ Jet<N> sfloor(const Jet<N>& x) { return Jet<N>(floorf(x.v)); }

// ---------------------------------------------------------------- vectors over S
template <class S>
struct T3 {
    S x, y, z;
};
template <class S> __device__ __forceinline__ T3<S> t3(const S& a, const S& b, const S& c) { return T3<S>{a, b, c}; }
template <class S> __device__ __forceinline__ T3<S> t3c(V3 v) { return T3<S>{S(v.x), S(v.y), S(v.z)}; }
template <class S> __device__ __forceinline__ T3<S> operator+(const T3<S>& a, const T3<S>& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class S> __device__ __forceinline__ T3<S> operator-(const T3<S>& a, const T3<S>& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class S> __device__ __forceinline__ T3<S> operator-(const T3<S>& a) { return {-a.x, -a.y, -a.z}; }
template <class S> __device__ __forceinline__ T3<S> operator*(const T3<S>& a, const T3<S>& b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class S> __device__ __forceinline__ T3<S> operator*(const T3<S>& a, const S& s) { return {a.x * s, a.y * s, a.z * s}; }
template <class S> __device__ __forceinline__ T3<S> operator*(const S& s, const T3<S>& a) { return {a.x * s, a.y * s, a.z * s}; }
template <class S> __device__ __forceinline__ S tdot(const T3<S>& a, const T3<S>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class S> __device__ __forceinline__ T3<S> tconst(float a) { return T3<S>{S(a), S(a), S(a)}; }

// ---------------------------------------------------------------- forward expressions over S
// Each mirrors the float function of the same name in trace_path.hpp, operation for operation.
template <class S> __device__ __forceinline__ S sat(const S& v) { return smin(smax(v, S(0.f)), S(1.f)); }
template <class S> __device__ __forceinline__ T3<S> sat3(const T3<S>& v) { return {sat(v.x), sat(v.y), sat(v.z)}; }
template <class S> __device__ __forceinline__ S clamp_rough(const S& r) { return smin(smax(r, S(kMinRough)), S(1.f)); }
template <class S> __device__ __forceinline__ T3<S> nonneg(const T3<S>& v) {
    return {smax(v.x, S(0.f)), smax(v.y, S(0.f)), smax(v.z, S(0.f))};
}
template <class S> __device__ __forceinline__ T3<S> fsn(const T3<S>& v, const T3<S>& fb) {
    const S l2 = tdot(v, v);
    return val(l2) > kBrdfEps ? v * srsqrt(l2) : fb;
}
template <class S> __device__ __forceinline__ T3<S> safe_normalize(const T3<S>& v) {
    const S l = tdot(v, v);
    return val(l) > 0.f ? v * srsqrt(l) : v;
}
template <class S> __device__ __forceinline__ void onb(const T3<S>& n, T3<S>& b1, T3<S>& b2) {
    const float sign = copysignf(1.0f, val(n.z));
    const S a = S(-1.0f) / (S(sign) + n.z);
    const S b = n.x * n.y * a;
    b1 = t3(S(1.0f) + sign * n.x * n.x * a, sign * b, -sign * n.x);
    b2 = t3(b, S(sign) + n.y * n.y * a, -n.y);
}
template <class S> __device__ __forceinline__ T3<S> to_world(const T3<S>& n, const T3<S>& l) {
    T3<S> t, b;
    onb(n, t, b);
    return t * l.x + b * l.y + n * l.z;
}
template <class S> __device__ __forceinline__ T3<S> sample_diffuse(const T3<S>& n, float u1, float u2) {
    const float xi = sat(u1), r = sqrtf(xi), phi = 2.0f * kPiF * sat(u2);
    float s, c;
    sincosf(phi, &s, &c);
    const float z = sqrtf(fmaxf(0.f, 1.f - xi));
    return to_world(n, t3(S(r * c), S(r * s), S(z)));
}
template <class S> __device__ __forceinline__ T3<S> sample_ggx_h(const T3<S>& n, const S& rough, float u1, float u2) {
    const S a = rough * rough, a2 = a * a;
    const float xi = sat(u1);
    const S den = smax((a2 - 1.f) * xi + 1.f, S(kBrdfEps));
    const S ct = ssqrt(smax(S(0.f), (1.f - xi) / den));
    const S st = ssqrt(smax(S(0.f), 1.f - ct * ct));
    float s, c;
    sincosf(2.0f * kPiF * sat(u2), &s, &c);
    return to_world(n, t3(st * c, st * s, ct));
}

// eval_ls<false> of trace_path.hpp: BRDF value (times cosine) and sampling pdf
template <class S>
__device__ __forceinline__ T3<S> eval_ls(const T3<S>& wo, const T3<S>& n, const T3<S>& albedo, const S& metal,
                                         const S& roughness, const T3<S>& ldir, S& pdf) {
    const T3<S> L = fsn(ldir, n);
    const S NV = sat(tdot(n, wo)), NL = sat(tdot(n, L));
    pdf = S(0.f);
    if (val(NV) <= 0.f || val(NL) <= 0.f) return tconst<S>(0.f);
    const S rough = clamp_rough(roughness);
    const T3<S> f0 = tconst<S>(0.04f) + (albedo - tconst<S>(0.04f)) * metal;
    const T3<S> H = fsn(wo + L, n);
    const S NH = sat(tdot(n, H)), VH = sat(tdot(wo, H));
    const S x = 1.f - sat(VH), x2 = x * x, q = x2 * x2 * x;
    const T3<S> F = f0 + (tconst<S>(1.f) - f0) * q;
    const float invPi = 1.0f / kPiF;
    const T3<S> diffCol = albedo * (1.f - metal);
    const S diffScale = NL * invPi;
    const T3<S> diffuse = diffCol * (tconst<S>(1.f) - F) * diffScale;
    const S al = rough * rough, a2 = al * al;
    const S dDen = smax(NH * NH * (a2 - 1.f) + 1.f, S(kBrdfEps));
    const S D = a2 * invPi / smax(dDen * dDen, S(kBrdfEps));
    const S k = 0.5f * rough * rough;
    const S Dv = smax(NV * (1.f - k) + k, S(kBrdfEps));
    const S Dl = smax(NL * (1.f - k) + k, S(kBrdfEps));
    const S Gv = NV / Dv, Gl = NL / Dl, G = Gv * Gl;
    const S specScale = D * G / smax(4.f * NV, S(kBrdfEps));
    const T3<S> spec = F * specScale;
    const S specPdf = val(VH) > 0.f ? D * NH / smax(4.f * VH, S(kBrdfEps)) : S(0.f);
    pdf = 0.5f * (diffScale + specPdf);
    return nonneg(diffuse + spec);
}

// sample_throughput<false>: value = f cos / pdf of the sampled direction L
template <class S>
__device__ __forceinline__ T3<S> sample_throughput(const T3<S>& wo, const T3<S>& n, const T3<S>& albedo, const S& metal,
                                                   const S& roughness, V3 rnd, T3<S>& L, S& pdf) {
    const S rough = clamp_rough(roughness);
    L = n;
    pdf = S(0.f);
    if (rnd.z < 0.5f) {
        L = sample_diffuse(n, rnd.x, rnd.y);
        if (!(val(tdot(n, L)) > 0.f)) return tconst<S>(0.f);
    } else {
        const T3<S> H = sample_ggx_h(n, rough, rnd.x, rnd.y);
        L = H * (2.f * tdot(wo, H)) - wo;
        if (!(val(tdot(n, L)) > 0.f)) return tconst<S>(0.f);
    }
    S p;
    const T3<S> bc = eval_ls(wo, n, albedo, metal, roughness, L, p);
    pdf = p;
    if (val(p) <= kBrdfEps) {
        pdf = S(0.f);
        return tconst<S>(0.f);
    }
    const S invPdf = S(1.f) / smax(p, S(kBrdfEps));
    return nonneg(bc * invPdf);
}

template <class S>
__device__ __forceinline__ void shading_frame(const T3<S>& d, const T3<S>& sn, const T3<S>& alb_raw, const S& met_raw,
                                              bool metallic, T3<S>& n, T3<S>& wo, T3<S>& alb, S& met) {
    n = fsn(sn, t3(S(0.f), S(0.f), S(1.f)));
    wo = fsn(-d, n);
    alb = sat3(alb_raw);
    met = metallic ? sat(met_raw) : S(0.f);
}

// environment lookup of a (world) direction: environment.cuh:123-158, 358-366, 49-87
template <class S>
__device__ __forceinline__ T3<S> background(const DeviceScene& Sc, const T3<S>& rd) {
    float sz, cz;
    sincosf(0.5f * kPiF, &sz, &cz);
    const T3<S> dir = t3(rd.x * cz - rd.y * sz, rd.x * sz + rd.y * cz, rd.z);
    const S theta = satan2(dir.x, dir.y);
    const S zcl = smin(S(1.f), smax(S(-1.f), -dir.z));
    const S phi = sasin(zcl);
    S u = (theta + kPiF) * (0.5f * (1.0f / kPiF));
    S v = 0.5f + phi * (1.0f / kPiF);
    u = u - sfloor(u);
    v = smin(smax(v, S(0.f)), S(1.f));
    const S x = u * float(Sc.env_w) - 0.5f, y = v * float(Sc.env_h) - 0.5f;
    const int x0 = int(floorf(val(x))), y0 = int(floorf(val(y)));
    const S ax = x - float(x0), ay = y - float(y0);
    auto wrap = [](int q, int s) { int w = q % s; return w < 0 ? w + s : w; };
    const int xa = wrap(x0, Sc.env_w), xb = wrap(x0 + 1, Sc.env_w);
    const int ya = max(0, min(y0, Sc.env_h - 1)), yb = max(0, min(y0 + 1, Sc.env_h - 1));
    const S w00 = (1.f - ax) * (1.f - ay), w10 = ax * (1.f - ay), w01 = (1.f - ax) * ay, w11 = ax * ay;
    const V3 a = xyz(Sc.env[ya * Sc.env_w + xa]), b = xyz(Sc.env[ya * Sc.env_w + xb]);
    const V3 c = xyz(Sc.env[yb * Sc.env_w + xa]), e = xyz(Sc.env[yb * Sc.env_w + xb]);
    return t3(w00 * a.x + w10 * b.x + w01 * c.x + w11 * e.x, w00 * a.y + w10 * b.y + w01 * c.y + w11 * e.y,
              w00 * a.z + w10 * b.z + w01 * c.z + w11 * e.z);
}

// sh_radiance of trace_path.hpp (gaussianParticles.cuh:48-87) in the ray direction
template <class S>
__device__ __forceinline__ T3<S> sh_radiance(const DeviceScene& Sc, uint32_t id, const T3<S>& d) {
    const float* c = Sc.sh + size_t(Sc.sh_stride) * id;
    auto C = [&](int k) { return v3(c[3 * k], c[3 * k + 1], c[3 * k + 2]); };
    auto acc = [](T3<S>& r, const S& w, V3 k) { r = r + t3(w * k.x, w * k.y, w * k.z); };
    T3<S> r = t3c<S>(0.28209479177387814f * C(0));
    const int deg = Sc.sph_degree;
    if (deg > 0) {
        const S x = d.x, y = d.y, z = d.z;
        const float C1 = 0.4886025119029199f;
        acc(r, -(C1 * y), C(1));
        acc(r, C1 * z, C(2));
        acc(r, -(C1 * x), C(3));
        if (deg > 1) {
            const S xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
            acc(r, 1.0925484305920792f * xy, C(4));
            acc(r, -1.0925484305920792f * yz, C(5));
            acc(r, 0.31539156525252005f * (2.f * zz - xx - yy), C(6));
            acc(r, -1.0925484305920792f * xz, C(7));
            acc(r, 0.5462742152960396f * (xx - yy), C(8));
            if (deg > 2) {
                acc(r, -0.5900435899266435f * y * (3.f * xx - yy), C(9));
                acc(r, 2.890611442640554f * xy * z, C(10));
                acc(r, -0.4570457994644658f * y * (4.f * zz - xx - yy), C(11));
                acc(r, 0.3731763325901154f * z * (2.f * zz - 3.f * xx - 3.f * yy), C(12));
                acc(r, -0.4570457994644658f * x * (4.f * zz - xx - yy), C(13));
                acc(r, 1.445305721320277f * z * (xx - yy), C(14));
                acc(r, -0.5900435899266435f * x * (xx - 3.f * yy), C(15));
            }
        }
    }
    r = r + tconst<S>(0.5f);
    return nonneg(r);
}

// SH basis values Y_k(d) as sh_radiance weighs coefficient k (constants and signs included)
__device__ __forceinline__ int sh_basis(int deg, V3 d, float* Y) {
    Y[0] = 0.28209479177387814f;
    if (deg == 0) return 1;
    const float x = d.x, y = d.y, z = d.z, C1 = 0.4886025119029199f;
    Y[1] = -(C1 * y);
    Y[2] = C1 * z;
    Y[3] = -(C1 * x);
    if (deg == 1) return 4;
    const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
    Y[4] = 1.0925484305920792f * xy;
    Y[5] = -1.0925484305920792f * yz;
    Y[6] = 0.31539156525252005f * (2.f * zz - xx - yy);
    Y[7] = -1.0925484305920792f * xz;
    Y[8] = 0.5462742152960396f * (xx - yy);
    if (deg == 2) return 9;
    Y[9] = -0.5900435899266435f * y * (3.f * xx - yy);
    Y[10] = 2.890611442640554f * xy * z;
    Y[11] = -0.4570457994644658f * y * (4.f * zz - xx - yy);
    Y[12] = 0.3731763325901154f * z * (2.f * zz - 3.f * xx - 3.f * yy);
    Y[13] = -0.4570457994644658f * x * (4.f * zz - xx - yy);
    Y[14] = 1.445305721320277f * z * (xx - yy);
    Y[15] = -0.5900435899266435f * x * (xx - 3.f * yy);
    return 16;
}

// ---------------------------------------------------------------- local Jacobians by jets
// Vertex block: the BRDF sample leaving a vertex (pathTracer.cuh:430-459) as a function of
// the segment direction d (through wo), the normalised shading normal, albedo, roughness
// and metallic of the vertex.  Outputs: sampled direction L (the next segment's direction),
// its unit vector nL (which offsets the next origin), the throughput factor and the pdf.
constexpr int kVtxIn = 11;  // d[3], sn[3], alb[3], rough, metal

template <class S>
__device__ __forceinline__ void vertex_block(const S* x, V3 u, bool metallic, T3<S>& L, T3<S>& nL, T3<S>& value, S& pdf) {
    T3<S> n, wo, alb;
    S met;
    shading_frame(t3(x[0], x[1], x[2]), t3(x[3], x[4], x[5]), t3(x[6], x[7], x[8]), x[10], metallic, n, wo, alb, met);
    value = sample_throughput(wo, n, alb, met, clamp_rough(x[9]), u, L, pdf);
    nL = safe_normalize(L);
}

template <int N, int BASE>
__device__ __forceinline__ void vertex_pass(const float* x, V3 u, bool metallic, V3 gL, V3 gnL, V3 gv, float gp, float* gx) {
    Jet<N> in[kVtxIn];
#pragma unroll
    for (int i = 0; i < kVtxIn; ++i) {
        in[i] = Jet<N>(x[i]);
        if (i >= BASE && i < BASE + N) in[i].d[i - BASE] = 1.f;
    }
    T3<Jet<N>> L, nL, value;
    Jet<N> pdf;
    vertex_block(in, u, metallic, L, nL, value, pdf);
#pragma unroll
    for (int j = 0; j < N; ++j)
        if (BASE + j < kVtxIn)
            gx[BASE + j] = gL.x * L.x.d[j] + gL.y * L.y.d[j] + gL.z * L.z.d[j] + gnL.x * nL.x.d[j] + gnL.y * nL.y.d[j] +
                           gnL.z * nL.z.d[j] + gv.x * value.x.d[j] + gv.y * value.y.d[j] + gv.z * value.z.d[j] +
                           gp * pdf.d[j];
}

// NEE block: BRDF value and pdf toward a fixed light direction wi (pathTracer.cuh:461-504)
// as a function of the vertex's normalised shading normal, albedo, roughness and metallic.
constexpr int kNeeIn = 8;  // sn[3], alb[3], rough, metal

template <int N, int BASE>
__device__ __forceinline__ void nee_pass(const float* x, V3 d, V3 wi, bool metallic, V3 gv, float gp, float* gx) {
    Jet<N> in[kNeeIn];
#pragma unroll
    for (int i = 0; i < kNeeIn; ++i) {
        in[i] = Jet<N>(x[i]);
        if (i >= BASE && i < BASE + N) in[i].d[i - BASE] = 1.f;
    }
    T3<Jet<N>> n, wo, alb;
    Jet<N> met, pdf;
    shading_frame(t3c<Jet<N>>(d), t3(in[0], in[1], in[2]), t3(in[3], in[4], in[5]), in[7], metallic, n, wo, alb, met);
    const T3<Jet<N>> value = eval_ls(wo, n, alb, met, clamp_rough(in[6]), t3c<Jet<N>>(wi), pdf);
#pragma unroll
    for (int j = 0; j < N; ++j)
        if (BASE + j < kNeeIn)
            gx[BASE + j] = gv.x * value.x.d[j] + gv.y * value.y.d[j] + gv.z * value.z.d[j] + gp * pdf.d[j];
}

__device__ __forceinline__ T3<Jet<3>> dir_jet(V3 d) {
    T3<Jet<3>> r = t3c<Jet<3>>(d);
    r.x.d[0] = 1.f;
    r.y.d[1] = 1.f;
    r.z.d[2] = 1.f;
    return r;
}
__device__ __forceinline__ V3 contract3(const T3<Jet<3>>& o, V3 g) {
    return v3(g.x * o.x.d[0] + g.y * o.y.d[0] + g.z * o.z.d[0], g.x * o.x.d[1] + g.y * o.y.d[1] + g.z * o.z.d[1],
              g.x * o.x.d[2] + g.y * o.y.d[2] + g.z * o.z.d[2]);
}

// ---------------------------------------------------------------- scalar helpers
__device__ __forceinline__ float dmis_da(float a, float b) {  // d/da of a^2 / max(a^2 + b^2, 1e-6)
    const float den = a * a + b * b;
    return den > 1e-6f ? 2.f * a * b * b / (den * den) : 2.f * a / 1e-6f;
}
__device__ __forceinline__ float dmis_db(float a, float b) {
    const float den = a * a + b * b;
    return den > 1e-6f ? -2.f * a * a * b / (den * den) : 0.f;
}
__device__ __forceinline__ bool nonzero(V3 v) { return v.x != 0.f || v.y != 0.f || v.z != 0.f; }
__device__ __forceinline__ V3 load3(const float* p, size_t i) { return v3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }

// d response(degree, d2) / d d2, given gres = response(degree, d2)
__device__ __forceinline__ float dresponse(int degree, float d2, float gres) {
    switch (degree) {
    case 8: return gres * (-0.000685871056241f * 4.f) * d2 * d2 * d2;
    case 5: return gres * (-0.0185185185185f * 2.5f) * d2 * sqrtf(d2);
    case 4: return gres * (-0.0555555555556f * 2.f) * d2;
    case 3: return gres * (-0.166666666667f * 1.5f) * sqrtf(d2);
    case 1: return d2 > 0.f ? gres * (-0.75f) * rsqrtf(d2) : 0.f;
    case 0: return (d2 > 0.f && gres > 0.f) ? -0.5f * 0.329630334487f * rsqrtf(d2) : 0.f;
    default: return -0.5f * gres;
    }
}

// env texel gradients of a bilinear lookup in direction rd (the 4 footprint texels)
__device__ __forceinline__ void env_texels_adj(const DeviceScene& S, V3 rd, V3 g, float* denv) {
    const V3 dir = env_rotate(rd);
    float u, v;
    equirect_uv(dir, u, v);
    const Foot f = footprint(S, u, v);
    auto add = [&](int x, int y, float w) {
        float* p = denv + 4 * (size_t(y) * S.env_w + x);
        atomicAdd(p + 0, w * g.x);
        atomicAdd(p + 1, w * g.y);
        atomicAdd(p + 2, w * g.z);
    };
    add(f.x0, f.y0, f.w00);
    add(f.x1, f.y0, f.w10);
    add(f.x0, f.y1, f.w01);
    add(f.x1, f.y1, f.w11);
}

// ---------------------------------------------------------------- per-hit geometry adjoint
// Gaussian id's alpha and hit distance along ray r as functions of its position mu, rotation
// quaternion q (w,x,y,z, used as given: R(q) of gaussianParticles.cuh:26-46), scale s and
// density sigma and of the ray:
//   M = diag(1/s) R^T, gro = M (o - mu), grd = M d / |M d|, gd = |grd x gro|^2,
//   alpha = min(0.99, response(gd) sigma), hitT = |s * grd * (grd . -gro)|.
// ga = dLoss/dalpha, gh = dLoss/dhitT.  Adds the parameter gradients into dg[0..10]
// (the [N,12] density row: pos, density, quat, scale) and the ray gradients into go, gdr.
__device__ __forceinline__ void hit_geom_adj(const DeviceScene& S, const float* prm, const Ray& r, float ga, float gh,
                                             float* dg, V3& go, V3& gdr) {
    const V3 mu = v3(prm[0], prm[1], prm[2]);
    const float sigma = prm[3];
    const float qw = prm[4], qx = prm[5], qy = prm[6], qz = prm[7];
    const V3 s = v3(prm[8], prm[9], prm[10]);
    float R[3][3];
    R[0][0] = 1.f - 2.f * (qy * qy + qz * qz); R[0][1] = 2.f * (qx * qy - qw * qz); R[0][2] = 2.f * (qx * qz + qw * qy);
    R[1][0] = 2.f * (qx * qy + qw * qz); R[1][1] = 1.f - 2.f * (qx * qx + qz * qz); R[1][2] = 2.f * (qy * qz - qw * qx);
    R[2][0] = 2.f * (qx * qz - qw * qy); R[2][1] = 2.f * (qy * qz + qw * qx); R[2][2] = 1.f - 2.f * (qx * qx + qy * qy);
    const float is[3] = {1.f / s.x, 1.f / s.y, 1.f / s.z};
    float M[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) M[a][c] = R[c][a] * is[a];
    const V3 p = r.o - mu;
    const V3 gro = v3(M[0][0] * p.x + M[0][1] * p.y + M[0][2] * p.z, M[1][0] * p.x + M[1][1] * p.y + M[1][2] * p.z,
                      M[2][0] * p.x + M[2][1] * p.y + M[2][2] * p.z);
    const V3 grdu = v3(M[0][0] * r.d.x + M[0][1] * r.d.y + M[0][2] * r.d.z, M[1][0] * r.d.x + M[1][1] * r.d.y + M[1][2] * r.d.z,
                       M[2][0] * r.d.x + M[2][1] * r.d.y + M[2][2] * r.d.z);
    const float l2 = dot(grdu, grdu);
    const float il = l2 > 0.f ? rsqrtf(l2) : 1.f;
    const V3 grd = grdu * il;
    const V3 gc = cross(grd, gro);
    const float gd = dot(gc, gc);
    const float gres = response(S.kernel_degree, gd);
    V3 g_gro = v3s(0.f), g_grd = v3s(0.f);
    float g_sigma = 0.f;
    V3 g_s = v3s(0.f);
    if (ga != 0.f && gres * sigma < 0.99f) {
        g_sigma = ga * gres;
        const float g_gd = ga * sigma * dresponse(S.kernel_degree, gd, gres);
        g_gro += cross(gc, grd) * (2.f * g_gd);
        g_grd += cross(gro, gc) * (2.f * g_gd);
    }
    if (gh != 0.f) {
        const float proj = -dot(grd, gro);
        const V3 grds = s * grd * proj;
        const float ht = sqrtf(dot(grds, grds));
        if (ht > 0.f) {
            const V3 gg = grds * (gh / ht);
            g_s += gg * grd * proj;
            g_grd += s * gg * proj;
            const float gproj = dot(gg, s * grd);
            g_grd -= gro * gproj;
            g_gro -= grd * gproj;
        }
    }
    // grd = grdu / |grdu|
    const V3 g_grdu = l2 > 0.f ? (g_grd - grd * dot(grd, g_grd)) * il : g_grd;
    // gro = M p, grdu = M d
    const V3 g_p = v3(M[0][0] * g_gro.x + M[1][0] * g_gro.y + M[2][0] * g_gro.z,
                      M[0][1] * g_gro.x + M[1][1] * g_gro.y + M[2][1] * g_gro.z,
                      M[0][2] * g_gro.x + M[1][2] * g_gro.y + M[2][2] * g_gro.z);
    go += g_p;
    gdr += v3(M[0][0] * g_grdu.x + M[1][0] * g_grdu.y + M[2][0] * g_grdu.z,
              M[0][1] * g_grdu.x + M[1][1] * g_grdu.y + M[2][1] * g_grdu.z,
              M[0][2] * g_grdu.x + M[1][2] * g_grdu.y + M[2][2] * g_grdu.z);
    const float gro_[3] = {g_gro.x, g_gro.y, g_gro.z}, gru_[3] = {g_grdu.x, g_grdu.y, g_grdu.z};
    const float p_[3] = {p.x, p.y, p.z}, d_[3] = {r.d.x, r.d.y, r.d.z};
    float gR[3][3];
    float gs_[3] = {g_s.x, g_s.y, g_s.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float gM = gro_[a] * p_[c] + gru_[a] * d_[c];
            gR[c][a] = gM * is[a];
            gs_[a] -= gM * M[a][c] * is[a];
        }
    }
    const float gw = 2.f * (-qz * gR[0][1] + qy * gR[0][2] + qz * gR[1][0] - qx * gR[1][2] - qy * gR[2][0] + qx * gR[2][1]);
    const float gx = 2.f * (qy * gR[0][1] + qz * gR[0][2] + qy * gR[1][0] - 2.f * qx * gR[1][1] - qw * gR[1][2] +
                            qz * gR[2][0] + qw * gR[2][1] - 2.f * qx * gR[2][2]);
    const float gy = 2.f * (-2.f * qy * gR[0][0] + qx * gR[0][1] + qw * gR[0][2] + qx * gR[1][0] + qz * gR[1][2] -
                            qw * gR[2][0] + qz * gR[2][1] - 2.f * qy * gR[2][2]);
    const float gz = 2.f * (-2.f * qz * gR[0][0] - qw * gR[0][1] + qx * gR[0][2] + qw * gR[1][0] - 2.f * qz * gR[1][1] +
                            qy * gR[1][2] + qx * gR[2][0] + qy * gR[2][1]);
    atomicAdd(dg + 0, -g_p.x);
    atomicAdd(dg + 1, -g_p.y);
    atomicAdd(dg + 2, -g_p.z);
    atomicAdd(dg + 3, g_sigma);
    atomicAdd(dg + 4, gw);
    atomicAdd(dg + 5, gx);
    atomicAdd(dg + 6, gy);
    atomicAdd(dg + 7, gz);
    atomicAdd(dg + 8, gs_[0]);
    atomicAdd(dg + 9, gs_[1]);
    atomicAdd(dg + 10, gs_[2]);
}

// ---------------------------------------------------------------- vertex records
// Per path vertex k, 7 float4 in structure-of-arrays order (record word w of vertex k of
// path pix at scratch[(k * 7 + w) * paths + pix]):
//   0 (o.xyz, d.x)  1 (d.yz, T, depth)  2 (sn.xyz, alb.x)  3 (alb.yz, rough, metal)
//   4 (radiance.xyz, flags)  5 (throughput.xyz, scatterPdf)  6 (lightPdf, u0, u1, u2)
constexpr int kVtxWords = 7;
constexpr uint32_t kVIValid = 1u, kVValid = 2u;

}  // namespace adj

// float4 words of the vertex records of a launch of `paths` paths
__host__ __device__ inline uint64_t adjoint_scratch_words(uint64_t paths, uint32_t max_bounces) {
    return paths * uint64_t(max_bounces > 0 ? max_bounces : 1) * adj::kVtxWords;
}

namespace adj {

struct Vtx {
    V3 o, d, sn, alb, rad, thr, u;
    float T, depth, rough, metal, spdf, lpdf;
    uint32_t flags;
};

__device__ __forceinline__ void put_vtx(float4* sc, size_t paths, size_t pix, int k, const Seg& c, V3 thr, V3 u) {
    float4* q = sc + size_t(k) * kVtxWords * paths + pix;
    const uint32_t fl = (c.ivalid ? kVIValid : 0u) | (c.valid ? kVValid : 0u);
    q[0 * paths] = make_float4(c.ray.o.x, c.ray.o.y, c.ray.o.z, c.ray.d.x);
    q[1 * paths] = make_float4(c.ray.d.y, c.ray.d.z, c.T, c.hitDist);
    q[2 * paths] = make_float4(c.sn.x, c.sn.y, c.sn.z, c.alb.x);
    q[3 * paths] = make_float4(c.alb.y, c.alb.z, c.rough, c.metal);
    q[4 * paths] = make_float4(c.radiance.x, c.radiance.y, c.radiance.z, __uint_as_float(fl));
    q[5 * paths] = make_float4(thr.x, thr.y, thr.z, c.scatterPdf);
    q[6 * paths] = make_float4(c.lightPdf, u.x, u.y, u.z);
}

__device__ __forceinline__ Vtx get_vtx(const float4* sc, size_t paths, size_t pix, int k) {
    const float4* q = sc + size_t(k) * kVtxWords * paths + pix;
    const float4 a = q[0], b = q[paths], c = q[2 * paths], d = q[3 * paths], e = q[4 * paths], f = q[5 * paths],
                 g = q[6 * paths];
    Vtx v;
    v.o = v3(a.x, a.y, a.z);
    v.d = v3(a.w, b.x, b.y);
    v.T = b.z;
    v.depth = b.w;
    v.sn = v3(c.x, c.y, c.z);
    v.alb = v3(c.w, d.x, d.y);
    v.rough = d.z;
    v.metal = d.w;
    v.rad = v3(e.x, e.y, e.z);
    v.flags = __float_as_uint(e.w);
    v.thr = v3(f.x, f.y, f.z);
    v.spdf = f.w;
    v.lpdf = g.x;
    v.u = v3(g.y, g.z, g.w);
    return v;
}

}  // namespace adj

namespace adj {

// Adjoint of one segment (pathTracer.cuh:216-311 as traced by ray_intersect): re-traces its
// hits front to back.  Segment totals A (radiance, material and shading-normal sums, depth)
// and T have adjoints gA and gT; total = gA . A.  For hit i with weight w_i = alpha_i T_i and
// per-hit feature adjoint G_i = gA . f_i,
//   dLoss/dalpha_i = G_i T_i - (sum_{j>i} G_j w_j + gT T) / (1 - alpha_i),
// where the suffix sum is total minus the prefix so far.
struct SegAdj {
    V3 gRad, gAlb, gSn;
    float gRough, gMet, gDepth, gT, Tfin, total;
};

template <int K, int SD>
__device__ __forceinline__ void seg_adjoint(const DeviceScene& S, const AdjIO& io, const Ray ray, bool selfOcc,
                                            bool dir_grad, const SegAdj& a, const Lane& L, V3& g_o, V3& g_d) {
    float t0, t1;
    scene_interval(S, ray, t0, t1);
    float lastHit = fmaxf(0.f, t0 - kEpsT);
    float T = 1.f, prefix = 0.f;
    const float minT = S.min_transmittance;
    const float dlen = sqrtf(dot(ray.d, ray.d));
    const bool want_rad = nonzero(a.gRad);
    const bool want_mat = nonzero(a.gAlb) || a.gRough != 0.f || a.gMet != 0.f;
    const bool want_sn = S.enable_normals && S.has_snormal && nonzero(a.gSn);
    float Y[16];
    int ncoef = 0;
    if (want_rad) ncoef = sh_basis(S.sph_degree, ray.d, Y);
    while (lastHit <= t1 && T > minT) {
        const int cnt = collect<K, SD>(S, ray, lastHit + kEpsT, t1 + kEpsT, L);
        if (cnt == 0) break;
        for (int k = 0; k < cnt; ++k) {
            if (!(T > minT)) break;
            const float t = __uint_as_float(L.at(k));
            const int slot = int(L.at(K + k));
            if (selfOcc && self_occluded(S, slot, t, ray.d)) {
                lastHit = fmaxf(lastHit, t);
                continue;
            }
            HitGeom h;
            const float al = hit_alpha(S, slot, ray, h);
            if (al > 0.f) {
                const uint32_t id = orig_id(S, slot);
                const float4 m0 = S.srec[2 * slot], m1 = S.srec[2 * slot + 1];
                const float hitT = t * dlen;
                const V3 snv = v3(m1.y, m1.z, m1.w);
                V3 c = v3s(0.f);
                if (want_rad) c = sh_radiance(S, id, ray.d);
                float G = dot(a.gRad, c) + dot(a.gAlb, xyz(m0)) + a.gRough * m0.w + a.gMet * m1.x + a.gDepth * hitT;
                if (S.enable_normals) G += dot(a.gSn, snv);
                const float w = al * T;
                prefix += G * w;
                const float ga = G * T - (a.total - prefix + a.gT * a.Tfin) / (1.f - al);
                if (want_mat) {
                    float* q = io.d_material + 5 * size_t(id);
                    atomicAdd(q + 0, w * a.gAlb.x);
                    atomicAdd(q + 1, w * a.gAlb.y);
                    atomicAdd(q + 2, w * a.gAlb.z);
                    atomicAdd(q + 3, w * a.gRough);
                    atomicAdd(q + 4, w * a.gMet);
                }
                if (want_sn) {
                    float* q = io.d_snormal + 3 * size_t(id);
                    atomicAdd(q + 0, w * a.gSn.x);
                    atomicAdd(q + 1, w * a.gSn.y);
                    atomicAdd(q + 2, w * a.gSn.z);
                }
                if (want_rad) {
                    // c = max(0, sum_k Y_k coef_k + 0.5) per channel: the clamp passes where c > 0
                    const V3 gw = v3(c.x > 0.f ? w * a.gRad.x : 0.f, c.y > 0.f ? w * a.gRad.y : 0.f,
                                     c.z > 0.f ? w * a.gRad.z : 0.f);
                    float* q = io.d_sh + size_t(S.sh_stride) * id;
                    for (int j = 0; j < ncoef; ++j) {
                        if (gw.x != 0.f) atomicAdd(q + 3 * j + 0, gw.x * Y[j]);
                        if (gw.y != 0.f) atomicAdd(q + 3 * j + 1, gw.y * Y[j]);
                        if (gw.z != 0.f) atomicAdd(q + 3 * j + 2, gw.z * Y[j]);
                    }
                    if (dir_grad) g_d += contract3(sh_radiance(S, id, dir_jet(ray.d)), a.gRad * w);
                }
                const float gh = a.gDepth * w;
                if (ga != 0.f || gh != 0.f)
                    hit_geom_adj(S, io.density + 12 * size_t(id), ray, ga, gh, io.d_density + 12 * size_t(id), g_o, g_d);
                T *= (1.f - al);
            }
            lastHit = fmaxf(lastHit, t);
        }
    }
}

}  // namespace adj

// One path of the exact adjoint: replay, then reverse sweep (see the top of this file).
template <int K, int SD>
__device__ __forceinline__ void trace_adjoint_path(const DeviceScene& S, const AdjIO& io, const Lane& L, const PixelId& px) {
    using namespace adj;
    const size_t pix = px.pix;
    const size_t paths = size_t(S.B) * S.H * S.W;
    Rng rng;
    rng.seed(Rng::launch_seed(uint32_t(px.x), uint32_t(px.y), uint32_t(px.z), S.frame));
    Path p;
    path_init(p);
    SegTape none;
    none.hits = nullptr;
    // ---- 1. replay (the forward's loop of referenceOptix.cu:27-71, k_trace_fwd)
    // Record k is written once segment k's vertex is done: with the uniforms of its BRDF
    // sample, or with zeros where the path ends there.  n = segments traced.
    ray_intersect<K, SD>(S, primary_ray(S, pix), p.cur, rng, false, true, true, nullptr, L, none);
    int n = 1;
    if (!p.cur.ivalid || S.max_bounces == 0u) {
        put_vtx(io.scratch, paths, pix, 0, p.cur, p.thr, v3s(0.f));
    } else {
        if (S.enable_mis) sample_nee<K, SD>(S, p, rng, L);
        for (unsigned depth = 0; depth < S.max_bounces; ++depth) {
            const int k = int(depth);
            if (!(depth + 1u < S.max_bounces && p.cur.ivalid)) {  // the forward's loop ends at this vertex
                put_vtx(io.scratch, paths, pix, k, p.cur, p.thr, v3s(0.f));
                break;
            }
            p.numBounces = depth + 1u;
            V3 nn, wo, alb;
            float met;
            shading_frame(S, p.cur, nn, wo, alb, met);
            const float r0 = rng.next_1d();
            const float r1 = rng.next_1d();
            const float r2 = rng.next_1d();
            V3 Ld = p.cur.ray.d;
            float sPdf = 0.f;
            const BG bt = sample_throughput<false>(wo, nn, alb, met, clamp_rough(p.cur.rough), v3(r0, r1, r2), Ld, sPdf,
                                                   S.enable_metallic);
            put_vtx(io.scratch, paths, pix, k, p.cur, p.thr, v3(r0, r1, r2));
            const V3 nt = p.thr * bt.value;
            if (is_zero(nt)) break;
            p.thr = nt;
            const Ray nr{p.cur.pos + safe_normalize(Ld) * kSelfOccOffset, Ld};
            const float lpdf = S.enable_mis ? env_pdf(S, Ld) : 0.f;
            p.cur.reset(nr, 0.f);
            if (S.enable_mis) {
                p.cur.scatterPdf = sPdf;
                p.cur.lightPdf = lpdf;
            }
            if (vmax(p.thr) < 1e-4f) break;
            if (!russian_roulette(S, p, rng)) break;
            ray_intersect<K, SD>(S, p.cur.ray, p.cur, rng, true, false, true, nullptr, L, none);
            n = k + 2;
        }
    }
    // ---- 2. reverse sweep
    const V3 dL = load3(io.pbr_grad, pix);
    const V3 dLnb = io.light_grad ? load3(io.light_grad, pix) : v3s(0.f);
    const V3 dDirect = io.comp_grad ? load3(io.comp_grad, 2 * pix) : v3s(0.f);
    const V3 dIndirect = io.comp_grad ? load3(io.comp_grad, 2 * pix + 1) : v3s(0.f);
    const bool mis_on = S.enable_mis != 0;
    const bool metallic = S.enable_metallic != 0;
    V3 g_thr = v3s(0.f), g_o = v3s(0.f), g_d = v3s(0.f);
    V3 gv_pos = v3s(0.f), gv_sn = v3s(0.f), gv_alb = v3s(0.f);
    float gv_rough = 0.f, gv_met = 0.f, g_spdf = 0.f;
    for (int k = n - 1; k >= 0; --k) {
        const Vtx R = get_vtx(io.scratch, paths, pix, k);
        const bool ivalid = (R.flags & kVIValid) != 0u, valid = (R.flags & kVValid) != 0u;
        SegAdj a;
        a.gRad = a.gAlb = a.gSn = v3s(0.f);
        a.gRough = a.gMet = a.gDepth = a.gT = 0.f;
        // contribution of segment k (accumulate, environment.cuh:636-715)
        if (k >= 1) {
            const bool misk = mis_on && k == 1;
            const float m = misk ? mis(R.spdf, R.lpdf) : 1.f;
            float g_m = 0.f;
            if (S.render_opts == 1 && ivalid) {
                const V3 gc = dL + dIndirect;
                g_thr += gc * R.rad * m;
                a.gRad = gc * R.thr * m;
                g_m = dot(gc, R.thr * R.rad);
            } else if (valid && has_env(S)) {
                const V3 bg = ptir::background(S, R.d);
                const V3 gc = dL + ((S.render_opts == 1 || k < 2) ? dDirect : dIndirect);
                g_thr += gc * bg * (R.T * m);
                a.gT += m * dot(gc, R.thr * bg);
                const V3 gbg = gc * R.thr * (R.T * m) + dLnb;
                env_texels_adj(S, R.d, gbg, io.d_env);
                g_d += contract3(background(S, dir_jet(R.d)), gbg);
                g_m = dot(gc, R.thr * bg) * R.T;
            }
            g_spdf = misk ? g_m * dmis_da(R.spdf, R.lpdf) : 0.f;
        }
        if (k == 0) {
            // G-buffer outputs of the primary segment
            if (io.rad_grad) a.gRad = load3(io.rad_grad, pix);
            if (io.dns_grad) a.gT -= io.dns_grad[pix];
            if (io.hit_grad) a.gDepth += io.hit_grad[2 * pix];
            if (ivalid) {
                if (io.mat_grad) {
                    const float* mg = io.mat_grad + 5 * pix;
                    gv_alb += v3(mg[0], mg[1], mg[2]);
                    gv_rough += mg[3];
                    gv_met += mg[4];
                }
                if (io.snrm_grad && S.enable_normals) gv_sn += load3(io.snrm_grad, pix);
            }
            // NEE at the primary vertex: counted by the forward once it reached segment 1
            if (mis_on && n >= 2 && p.elightPdf > 0.f) {
                const V3 wi = p.eray.d;
                const V3 Li = ptir::background(S, wi);
                const float lMis = mis(p.elightPdf, p.escatterPdf);
                const float inv = 1.f / fmaxf(p.elightPdf, 1e-6f);
                const V3 gc = dL + (S.render_opts == 0 ? dDirect : v3s(0.f));
                V3 nn, wo, alb;
                float met;
                Seg c0;
                c0.ray.d = R.d;
                c0.sn = R.sn;
                c0.alb = R.alb;
                c0.metal = R.metal;
                shading_frame(S, c0, nn, wo, alb, met);
                float sp;
                const BG bc = eval_ls<false>(wo, nn, alb, met, clamp_rough(R.rough), wi, sp, metallic);
                const V3 gval = gc * Li * (inv * lMis);
                const float gp = dot(gc, p.econtrib) * dmis_db(p.elightPdf, p.escatterPdf);
                env_texels_adj(S, wi, gc * bc.value * (inv * lMis), io.d_env);
                const float x[kNeeIn] = {R.sn.x, R.sn.y, R.sn.z, R.alb.x, R.alb.y, R.alb.z, R.rough, R.metal};
                float gx[kNeeIn];
                nee_pass<4, 0>(x, R.d, wi, metallic, gval, gp, gx);
                nee_pass<4, 4>(x, R.d, wi, metallic, gval, gp, gx);
                gv_sn += v3(gx[0], gx[1], gx[2]);
                gv_alb += v3(gx[3], gx[4], gx[5]);
                gv_rough += gx[6];
                gv_met += gx[7];
            }
        }
        // vertex quantities (interaction.h:47-68): pos = o + d depth / op, x = X_acc / op
        const float op = 1.f - R.T;
        if (ivalid) {
            const float inv = 1.f / fmaxf(op, 1e-12f);
            g_o += gv_pos;
            g_d += gv_pos * (R.depth * inv);
            const float gpd = dot(gv_pos, R.d);
            a.gDepth += gpd * inv;
            a.gSn = gv_sn * inv;
            a.gAlb = gv_alb * inv;
            a.gRough = gv_rough * inv;
            a.gMet = gv_met * inv;
            const float g_op = -inv * (gpd * R.depth * inv + dot(gv_sn, R.sn) + dot(gv_alb, R.alb) + gv_rough * R.rough +
                                       gv_met * R.metal);
            a.gT -= g_op;
        }
        const bool any = a.gT != 0.f || a.gDepth != 0.f || nonzero(a.gRad) || nonzero(a.gAlb) || a.gRough != 0.f ||
                         a.gMet != 0.f || nonzero(a.gSn);
        if (any) {
            a.Tfin = R.T;
            a.total = dot(a.gRad, R.rad) + op * (dot(a.gAlb, R.alb) + a.gRough * R.rough + a.gMet * R.metal) +
                      (S.enable_normals ? op * dot(a.gSn, R.sn) : 0.f) + a.gDepth * R.depth;
            seg_adjoint<K, SD>(S, io, Ray{R.o, R.d}, k > 0, k > 0, a, L, g_o, g_d);
        }
        if (k == 0) break;
        // the BRDF sample at vertex k-1 that produced segment k (pathTracer.cuh:430-459),
        // then Russian roulette
        const Vtx P = get_vtx(io.scratch, paths, pix, k - 1);
        const float x[kVtxIn] = {P.d.x, P.d.y, P.d.z, P.sn.x, P.sn.y, P.sn.z, P.alb.x, P.alb.y, P.alb.z, P.rough, P.metal};
        T3<float> Lf, nLf, vf;
        float pdf;
        vertex_block(x, P.u, metallic, Lf, nLf, vf, pdf);
        const V3 value = v3(vf.x, vf.y, vf.z);
        const V3 pre = P.thr * value;
        V3 g_pre = g_thr;
        if (S.rr_start != 0u && unsigned(k) >= S.rr_start) {
            const float q = fminf(1.0f, vmax(pre));
            g_pre = g_thr * (1.f / q);
            if (q < 1.f) {
                // q = max component (ties: x, then y, as fmaxf(x, fmaxf(y, z)) is differentiated)
                const float gq = -dot(g_thr, pre) / (q * q);
                if (!(pre.x < fmaxf(pre.y, pre.z))) g_pre.x += gq;
                else if (!(pre.y < pre.z)) g_pre.y += gq;
                else g_pre.z += gq;
            }
        }
        const V3 g_value = g_pre * P.thr;
        g_thr = g_pre * value;
        float gx[kVtxIn];
        vertex_pass<4, 0>(x, P.u, metallic, g_d, g_o * kSelfOccOffset, g_value, g_spdf, gx);
        vertex_pass<4, 4>(x, P.u, metallic, g_d, g_o * kSelfOccOffset, g_value, g_spdf, gx);
        vertex_pass<4, 8>(x, P.u, metallic, g_d, g_o * kSelfOccOffset, g_value, g_spdf, gx);
        gv_pos = g_o;
        gv_sn = v3(gx[3], gx[4], gx[5]);
        gv_alb = v3(gx[6], gx[7], gx[8]);
        gv_rough = gx[9];
        gv_met = gx[10];
        g_o = v3s(0.f);
        g_d = v3(gx[0], gx[1], gx[2]);
        g_spdf = 0.f;
    }
}

}  // namespace ptir
