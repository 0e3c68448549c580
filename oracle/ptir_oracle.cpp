// PTIR path-tracer ORACLE — test infrastructure only.
//
// A scalar CPU restatement, in double (or float) precision, of the reference's
// differentiable Gaussian path tracer `threedgptir_tracer`: its forward raygen
// (src/kernels/cuda/referenceOptix.cu:27-71) and its path-replay backward
// (src/kernels/cuda/referenceBwdOptix.cu:27-74) with the helpers in
// include/3dgptir/kernels/cuda/{pathTracer,gaussianParticles,material,
// environment,lightSampler,sampler}.cuh.  All paths below are relative to
// threedgptir_tracer/ in the reference unless they start with threedgrut/.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// this library, and only as the checker.  The product (ptir_amd/) never calls it.
//
// Pinning status (tests/test_oracle_pins.py, test_oracle_closed_forms.py,
// test_oracle_backward_fd.py):
//   pcg32, launch seed ........ pinned: published pcg32 KAT + closed form
//   particle response ......... pinned: closed form exp(-r^4/18), r=3 special case
//   compositing ............... pinned: single/disjoint Gaussian closed forms,
//                                brute-force vs BVH candidate sets
//   Gaussian geometry ......... pinned: rotated anisotropic Gaussians vs the quadratic
//                                form of Sigma^-1 (quaternion convention, peak, rho);
//                                primary normal vs the 3-sigma ellipsoid entry point
//                                (unrotated Gaussians; the rotated case is restated as
//                                written, gaussianParticles.cuh:433-435)
//   SH ........................ pinned: every coefficient vs scipy's complex SH in the
//                                3DGS real convention; orthonormality by quadrature
//   BRDF eval/sample/grads .... pinned: pdf integrates to 1, sample histogram,
//                                grads vs central differences of the value
//   env lookup/sampling ....... pinned: constant env, bilinear partition of unity,
//                                alias-table histogram vs pdf, integral of pdf = 1
//   full forward .............. pinned: closed-form plane scenes: a Lambert+GGX disc
//                                under a constant sky (NEE + MIS, one bounce), an
//                                INDIRECT floor under an emissive ceiling, a two-bounce
//                                floor + wall under a constant sky (direct and indirect)
//   ptir backward ............. pinned for albedo (MIS on and off), metallic and env
//                                texels (MIS off) by central differences of the FP64
//                                forward; linear in the output gradients.  Its
//                                roughness gradient (a forward-difference direction term
//                                by design) and the NEE-side terms the reference leaves
//                                disabled are restated as written: parity unpinned.
//
// Candidate set.  The reference gets candidates from OptiX: instance k is a unit cube
// under M_k = R_k diag(kscl_k) (src/particlePrimitives.cu:543-610) and the IS program
// (src/kernels/cuda/referenceOptix.cu:73-93, gaussianParticles.cuh:484-500) accepts
// t* = -o'.d'/d'.d' in (tmin, tmax) with |normalize(d') x o'|^2/|d'|^2 < 9.  OptiX's
// box culling depends on its traversal order, so this build fixes a deterministic rule:
// a Gaussian is a candidate iff the peak point o'+t* d' lies inside the unit cube, t* is
// in (tmin, tmax) and the distance test holds.  Hits are ordered by (t*, index).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

// ---------------------------------------------------------------- forward-mode dual number
// The oracle is instantiated on Dual to get exact directional derivatives of every output
// along a chosen direction in parameter space (the "exact adjoint" of the GPU is checked
// against them, and they against central differences).  Comparisons use the value only:
// discrete decisions are those of the FP64 evaluation, and the derivative is that of the
// smooth piece of the estimator the path lies on.
struct Dual {
    double v = 0.0, d = 0.0;
    Dual() = default;
    Dual(double x) : v(x), d(0.0) {}  // NOLINT: constants and inputs enter as duals with zero tangent
    Dual(double x, double t) : v(x), d(t) {}
    explicit operator double() const { return v; }
    explicit operator float() const { return float(v); }
    explicit operator int() const { return int(v); }
    Dual operator-() const { return {-v, -d}; }
    Dual& operator+=(const Dual& o) { v += o.v; d += o.d; return *this; }
    Dual& operator-=(const Dual& o) { v -= o.v; d -= o.d; return *this; }
    Dual& operator*=(const Dual& o) { d = d * o.v + v * o.d; v *= o.v; return *this; }
    Dual& operator/=(const Dual& o) { d = (d * o.v - v * o.d) / (o.v * o.v); v /= o.v; return *this; }
};
inline Dual operator+(Dual a, const Dual& b) { return a += b; }
inline Dual operator-(Dual a, const Dual& b) { return a -= b; }
inline Dual operator*(Dual a, const Dual& b) { return a *= b; }
inline Dual operator/(Dual a, const Dual& b) { return a /= b; }
inline Dual operator+(Dual a, double b) { return a += Dual(b); }
inline Dual operator-(Dual a, double b) { return a -= Dual(b); }
inline Dual operator*(Dual a, double b) { return {a.v * b, a.d * b}; }
inline Dual operator/(Dual a, double b) { return {a.v / b, a.d / b}; }
inline Dual operator+(double a, const Dual& b) { return Dual(a) + b; }
inline Dual operator-(double a, const Dual& b) { return Dual(a) - b; }
inline Dual operator*(double a, const Dual& b) { return {a * b.v, a * b.d}; }
inline Dual operator/(double a, const Dual& b) { return Dual(a) / b; }
inline bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
inline bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
inline bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
inline bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
inline bool operator==(const Dual& a, const Dual& b) { return a.v == b.v; }
inline bool operator!=(const Dual& a, const Dual& b) { return a.v != b.v; }
inline bool operator<(const Dual& a, double b) { return a.v < b; }
inline bool operator>(const Dual& a, double b) { return a.v > b; }
inline bool operator<=(const Dual& a, double b) { return a.v <= b; }
inline bool operator>=(const Dual& a, double b) { return a.v >= b; }
inline bool operator==(const Dual& a, double b) { return a.v == b; }
inline bool operator!=(const Dual& a, double b) { return a.v != b; }
inline bool operator<(double a, const Dual& b) { return a < b.v; }
inline bool operator>(double a, const Dual& b) { return a > b.v; }
inline double tangent_of(const Dual& x) { return x.d; }
inline double tangent_of(double) { return 0.0; }
inline double tangent_of(float) { return 0.0; }

namespace std {
inline Dual sqrt(const Dual& x) { const double r = std::sqrt(x.v); return {r, r > 0 ? x.d / (2 * r) : 0.0}; }
inline Dual exp(const Dual& x) { const double e = std::exp(x.v); return {e, e * x.d}; }
inline Dual log(const Dual& x) { return {std::log(x.v), x.d / x.v}; }
inline Dual sin(const Dual& x) { return {std::sin(x.v), std::cos(x.v) * x.d}; }
inline Dual cos(const Dual& x) { return {std::cos(x.v), -std::sin(x.v) * x.d}; }
inline Dual fabs(const Dual& x) { return x.v < 0 ? -x : x; }
inline Dual floor(const Dual& x) { return {std::floor(x.v), 0.0}; }
inline Dual round(const Dual& x) { return {std::round(x.v), 0.0}; }
inline Dual fmax(const Dual& a, const Dual& b) { return a.v >= b.v ? a : b; }
inline Dual fmin(const Dual& a, const Dual& b) { return a.v <= b.v ? a : b; }
inline Dual copysign(const Dual& a, const Dual& b) { return (a.v < 0) != (b.v < 0) ? -a : a; }
inline Dual atan2(const Dual& y, const Dual& x) {
    const double r2 = x.v * x.v + y.v * y.v;
    return {std::atan2(y.v, x.v), (x.v * y.d - y.v * x.d) / r2};
}
inline Dual asin(const Dual& x) { return {std::asin(x.v), x.d / std::sqrt(std::max(1.0 - x.v * x.v, 1e-300))}; }
inline Dual pow(const Dual& a, const Dual& b) {
    const double p = std::pow(a.v, b.v);
    return {p, p * (b.d * std::log(a.v) + b.v * a.d / a.v)};
}
template <> struct numeric_limits<Dual> {
    static Dual infinity() { return Dual(std::numeric_limits<double>::infinity()); }
};
}  // namespace std

namespace orc {

// Input value with its tangent (zero unless the input array has a tangent array).
template <class R> static inline R input(float v, const float* tangent, size_t i) {
    (void)tangent;
    (void)i;
    return R(v);
}
template <> inline Dual input<Dual>(float v, const float* tangent, size_t i) {
    return Dual(double(v), tangent ? double(tangent[i]) : 0.0);
}

// ---------------------------------------------------------------- vectors
template <class R>
struct V3 {
    R x, y, z;
};
template <class R> static inline V3<R> v3(R a, R b, R c) { return V3<R>{a, b, c}; }
template <class R> static inline V3<R> v3s(R a) { return V3<R>{a, a, a}; }
template <class R> static inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> static inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> static inline V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <class R> static inline V3<R> operator*(V3<R> a, V3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class R> static inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> static inline V3<R> operator*(R s, V3<R> a) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> static inline V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> static inline V3<R>& operator+=(V3<R>& a, V3<R> b) { a = a + b; return a; }
template <class R> static inline V3<R>& operator-=(V3<R>& a, V3<R> b) { a = a - b; return a; }
template <class R> static inline V3<R>& operator*=(V3<R>& a, V3<R> b) { a = a * b; return a; }
template <class R> static inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> static inline V3<R> cross(V3<R> a, V3<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> static inline bool is_zero(V3<R> a) { return a.x == 0 && a.y == 0 && a.z == 0; }
template <class R> static inline R vmax(V3<R> a) { return std::max(a.x, std::max(a.y, a.z)); }

// mathUtils.h:390 safe_normalize: v/|v| if |v|^2 > 0 else v.
template <class R> static inline V3<R> safe_normalize(V3<R> v) {
    const R l = dot(v, v);
    return l > R(0) ? v * (R(1) / std::sqrt(l)) : v;
}

// Duff et al. 2017 branchless orthonormal basis (mathUtils.h:442-448).
template <class R> static inline void onb(V3<R> n, V3<R>& b1, V3<R>& b2) {
    const R sign = std::copysign(R(1), n.z);
    const R a = R(-1) / (sign + n.z);
    const R b = n.x * n.y * a;
    b1 = v3<R>(R(1) + sign * n.x * n.x * a, sign * b, -sign * n.x);
    b2 = v3<R>(b, sign + n.y * n.y * a, -n.y);
}

// The reference adds 1e-9 to hit distances in float32 (pathTracer.cuh:221,239,326-335).
// Above 2^-5 that addition rounds back to the input, so a chunk that restarts at the last
// hit excludes it; below 2^-5 the distance advances.  This restates that rounding rule.
template <class R> static inline R adv(R x) { return x >= R(0.03125) ? x : x + R(1e-9); }
template <class R> static inline R back(R x) { return x >= R(0.03125) ? x : x - R(1e-9); }

// ---------------------------------------------------------------- PCG32 (sampler.cuh:26-84)
struct Pcg32 {
    uint64_t state = 0, inc = 1;
    static uint64_t launch_seed(uint32_t x, uint32_t y, uint32_t z, uint32_t frame) {
        uint64_t s = 1469598103934665603ull;
        s ^= uint64_t(x) + 1099511628211ull * s;
        s ^= uint64_t(y) + 1099511628211ull * s;
        s ^= uint64_t(z) + 1099511628211ull * s;
        s ^= uint64_t(frame) + 1099511628211ull * s;
        return s ? s : 1ull;
    }
    void seed(uint64_t s, uint64_t seq) {
        state = 0;
        inc = (seq << 1u) | 1ull;
        next_u32();
        state += s;
        next_u32();
    }
    uint32_t next_u32() {
        const uint64_t old = state;
        state = old * 6364136223846793005ull + inc;
        const uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = uint32_t(old >> 59u);
        return (xs >> rot) | (xs << ((-int32_t(rot)) & 31));
    }
    // low 24 bits scaled by 2^-24: exact in float and double.
    double next_1d() { return double(next_u32() & 0x00FFFFFFu) * (1.0 / 16777216.0); }
};

// ---------------------------------------------------------------- constants
constexpr double kPi = 3.14159265358979323846;
constexpr double kSelfOccOffset = 2e-2;                    // pathTracer.cuh:27
constexpr double kMaxSelfOcc = 1e-1 - kSelfOccOffset;       // pathTracer.cuh:28
constexpr double kSelfOccDot = 0.1;                          // pathTracer.cuh:29
constexpr double kBrdfEps = 1e-6;                            // material.cuh:25
constexpr double kMinRough = 0.05;                           // material.cuh:26
constexpr double kRoughGradEps = 1e-3;                       // material.cuh:29
constexpr double kEnvDirEps = 1e-3;                          // environment.cuh:16
constexpr double kInfDist = 1e20;                            // pathTracer.cuh:36
constexpr uint32_t kInvalid = 0xFFFFFFFFu;

}  // namespace orc

extern "C" {

// Inputs exactly as `OptixTracer::trace` receives them (src/optixTracer.cpp:909-930).
struct OracleScene {
    int32_t n;
    const float* density;    // [n,12] pos3, sigma, quat wxyz, scale3, pad
    const float* material;   // [n,5] albedo3, roughness, metallic
    const float* sh;         // [n, sh_stride]
    int32_t sh_stride;       // floats per Gaussian = 3*(max_deg+1)^2
    const float* snormal;    // [n,3] (may be null)
    int32_t env_h, env_w;    // env [env_h, env_w, 4] (0 = none)
    const float* env;
    int32_t tab_h, tab_w;    // alias table [3, tab_h, tab_w] (0 = none)
    const float* tab;
    float kernel_degree;
    float min_response;
    int32_t density_clamping;
    int32_t sph_degree;      // evaluated SH degree (n_active_features)
    int32_t enable_normals, enable_hitcounts, enable_mis, enable_metallic;
    int32_t render_opts;
    float min_transmittance;
    uint32_t max_bounces;
    uint32_t frame;
    int32_t kbuf;            // hits per chunk (16 in the reference)
    int32_t tab_kind;        // 0: alias table (prob, alias, pdf); 1: CDF table (conditional, marginal, pdf)
    uint32_t rr_start;       // Russian roulette from this path vertex on (0 = off); no reference counterpart
    // tangents of the inputs (optional, same shapes): the dual-number run differentiates
    // every output along this direction
    const float* t_density;
    const float* t_material;
    const float* t_sh;
    const float* t_snormal;
    const float* t_env;
};

struct OracleRays {
    int32_t B, H, W;
    const float* ray_to_world;  // 12 floats (first 3 rows of batch element 0)
    const float* ori;           // [B,H,W,3]
    const float* dir;           // [B,H,W,3]
    int64_t path_stride;        // evaluate every path_stride-th path (1 = all)
    int64_t path_offset;
    int64_t z_offset;           // added to the image index of the launch index (x, y, z)
};

// Outputs in double, [B,H,W,C] like the reference's 13 outputs (src/optixTracer.cpp:931-944).
struct OracleFwdOut {
    double *rad, *dns, *hit, *hit2, *distortion, *nrm, *snrm, *mat, *hits, *pbr, *light, *comp;
    double* visibility;  // [n]
    // [B,H,W] (optional): smallest relative distance of any discrete decision of the path to
    // its threshold (hit acceptance, candidate box and distance tests, transmittance cut-off,
    // interaction opacity, escape and shadow draws, self-occlusion, hemisphere tests, the
    // throughput cut-off, nearest env-table cell).  A path whose margin is well above float32
    // rounding takes the same decisions in any float32 evaluation of the same equations.
    double* margin;
};

struct OracleBwdIn {
    const double* pbr;        // [B,H,W,3] forward rayPbr
    const double* light;      // [B,H,W,3] forward rayLight
    const float* pbr_grad;    // [B,H,W,3]
    const float* light_grad;  // [B,H,W,3]
    const float* mat_grad;    // [B,H,W,5]
};

struct OracleBwdOut {
    double* dmat;  // [n,5]
    double* denv;  // [env_h, env_w, 4]
};

struct OracleCounters {
    int64_t paths, segments, candidates, accepted, accepted_primary, shadow_candidates, env_lookups;
    int64_t mat_atomics, env_atomics;
};

// 3DGRT primary-only backward (threedgrt_tracer/src/optixTracer.cpp:927-1012): the forward's
// outputs [B,H,W,C] and their gradients; hit_grad has one channel (the integrated depth).
struct OracleGrtIn {
    const double *rad, *dns, *hit, *hit2, *distortion, *snrm;
    const float *rad_grad, *dns_grad, *hit_grad, *hit2_grad, *distortion_grad, *snrm_grad;
    int32_t inclusive_end;  // 1: also process the hits at exactly the last forward hit distance
};

struct OracleGrtOut {
    double* ddensity;  // [n,12] position, density, quaternion, scale, pad
    double* dsh;       // [n, sh_stride]
    double* dsnormal;  // [n,3]
};
}

namespace orc {

template <class R>
struct Gauss {
    V3<R> mu, s, is;   // position, scale, 1/scale
    R rot[3][3];       // R (standard, column-vector convention)
    R sigma;
    V3<R> kscl;        // kernelScale(sigma) * s
    V3<R> lo, hi;      // world AABB of the instance cube
};

template <class R>
struct Mat {
    V3<R> albedo = v3s<R>(0);
    R rough = 0, metal = 0;
};

template <class R>
struct Inter {
    bool valid = false;
    V3<R> pos = v3s<R>(0), sn = v3s<R>(0);
    Mat<R> m, g;
};

template <class R>
struct Ray {
    V3<R> o = v3s<R>(0), d = v3s<R>(0);
};

template <class R>
struct Seg {  // payLoad.h:38-100 rayPayload
    Ray<R> ray;
    Inter<R> inter;
    V3<R> radiance = v3s<R>(0), light = v3s<R>(0), contribution = v3s<R>(0), normal = v3s<R>(0);
    V3<R> snRaw = v3s<R>(0);  // sum of weight x shading normal (the 3DGRT shading-normal output)
    R T = 1, hitDist = 0, hitDist2 = 0, lastHit = 0, distortion = 0, hits = 0;
    R scatterPdf = 0, lightPdf = 0;
    V3<R> dirGrad = v3s<R>(0);
    bool hit = false, valid = false;
    Seg() = default;
    Seg(const Ray<R>& r, R last) : ray(r), lastHit(last) {}
};

template <class R>
struct Pending {  // payLoad.h:102-140
    bool valid = false;
    Ray<R> ray;
    R opacity = 0, maxHit = 0;
    V3<R> dDir = v3s<R>(0);
    unsigned numBounces = 0;
};

template <class R>
struct Path {  // payLoad.h:142-168
    unsigned active = 1, numBounces = 0, maxBounces = 0;
    Seg<R> cur, emit;
    V3<R> L = v3s<R>(0), dL = v3s<R>(0), direct = v3s<R>(0), indirect = v3s<R>(0);
    V3<R> Lnb = v3s<R>(0), dLnb = v3s<R>(0), thr = v3s<R>(1);
    Pending<R> pend;
};

struct Cand {
    double t;
    uint32_t id;
};

// ---------------------------------------------------------------- scene
template <class R>
struct Scene {
    const OracleScene* in;
    std::vector<Gauss<R>> g;
    V3<R> lo, hi;  // scene AABB
    // median-split BVH over world AABBs
    struct Node {
        V3<R> lo, hi;
        int32_t left, right, first, count;
    };
    std::vector<Node> nodes;
    std::vector<uint32_t> order;
    int ncoef;
    bool has_env() const { return in->env && in->env_h > 0 && in->env_w > 0; }
    bool has_env_light() const { return has_env() && in->tab && in->tab_h > 0 && in->tab_w > 0; }
};

// particlePrimitives.cu:27-51 kernelScale
template <class R>
static R kernel_scale(R density, R min_response, bool clamping, R degree) {
    const R modulation = clamping ? density : R(1);
    const R minR = std::min(min_response / modulation, R(0.97));
    if (degree < 0) {
        const R k = std::fabs(degree);
        const R s = R(1) / std::pow(R(3), k);
        return std::pow((R(1) / (std::log(minR) - R(1)) + R(1)) / s, R(1) / k);
    }
    if (degree == 0) return ((R(1) - minR) / R(3)) / R(-0.329630334487);
    const R a = R(-4.5) / std::pow(R(3), degree);
    return std::pow(std::log(minR) / a, R(1) / degree);
}

// gaussianParticles.cuh:260-301 particleResponse (generalized Gaussian exp(-4.5 (r/3)^(2n)))
template <class R>
static R response(int degree, R d2) {
    switch (degree) {
    case 8: { const R q = d2 * d2; return std::exp(R(-4.5 / 6561.0) * q * q); }
    case 5: return std::exp(R(-4.5 / 243.0) * d2 * d2 * std::sqrt(d2));
    case 4: return std::exp(R(-4.5 / 81.0) * d2 * d2);
    case 3: return std::exp(R(-4.5 / 27.0) * d2 * std::sqrt(d2));
    case 1: return std::exp(R(-1.5) * std::sqrt(d2));
    case 0: return std::max(R(1) + R(-0.329630334487) * std::sqrt(d2), R(0));
    default: return std::exp(R(-0.5) * d2);
    }
}

template <class R>
static void build_scene(Scene<R>& sc, const OracleScene* in) {
    sc.in = in;
    sc.ncoef = (in->sph_degree + 1) * (in->sph_degree + 1);
    const int n = in->n;
    sc.g.resize(n);
    sc.lo = v3s<R>(std::numeric_limits<R>::infinity());
    sc.hi = v3s<R>(-std::numeric_limits<R>::infinity());
    for (int i = 0; i < n; ++i) {
        const float* p = in->density + 12 * i;
        auto P = [&](int k) { return input<R>(p[k], in->t_density, 12 * size_t(i) + k); };
        Gauss<R>& G = sc.g[i];
        G.mu = v3<R>(P(0), P(1), P(2));
        G.sigma = P(3);
        const R w = P(4), x = P(5), y = P(6), z = P(7);
        // gaussianParticles.cuh:26-46: standard rotation matrix of quaternion (w,x,y,z)
        G.rot[0][0] = 1 - 2 * (y * y + z * z); G.rot[0][1] = 2 * (x * y - w * z); G.rot[0][2] = 2 * (x * z + w * y);
        G.rot[1][0] = 2 * (x * y + w * z); G.rot[1][1] = 1 - 2 * (x * x + z * z); G.rot[1][2] = 2 * (y * z - w * x);
        G.rot[2][0] = 2 * (x * z - w * y); G.rot[2][1] = 2 * (y * z + w * x); G.rot[2][2] = 1 - 2 * (x * x + y * y);
        G.s = v3<R>(P(8), P(9), P(10));
        G.is = v3<R>(R(1) / G.s.x, R(1) / G.s.y, R(1) / G.s.z);
        const R ks = kernel_scale<R>(G.sigma, in->min_response, in->density_clamping != 0, in->kernel_degree);
        G.kscl = G.s * ks;
        // world AABB of R diag(kscl) [-1,1]^3 + mu
        V3<R> ext;
        ext.x = std::fabs(G.rot[0][0]) * G.kscl.x + std::fabs(G.rot[0][1]) * G.kscl.y + std::fabs(G.rot[0][2]) * G.kscl.z;
        ext.y = std::fabs(G.rot[1][0]) * G.kscl.x + std::fabs(G.rot[1][1]) * G.kscl.y + std::fabs(G.rot[1][2]) * G.kscl.z;
        ext.z = std::fabs(G.rot[2][0]) * G.kscl.x + std::fabs(G.rot[2][1]) * G.kscl.y + std::fabs(G.rot[2][2]) * G.kscl.z;
        G.lo = G.mu - ext;
        G.hi = G.mu + ext;
        sc.lo = v3<R>(std::min(sc.lo.x, G.lo.x), std::min(sc.lo.y, G.lo.y), std::min(sc.lo.z, G.lo.z));
        sc.hi = v3<R>(std::max(sc.hi.x, G.hi.x), std::max(sc.hi.y, G.hi.y), std::max(sc.hi.z, G.hi.z));
    }
    // median-split BVH, leaves of <= 4
    sc.order.resize(n);
    for (int i = 0; i < n; ++i) sc.order[i] = i;
    sc.nodes.clear();
    if (n == 0) return;
    struct Item { int node, first, count; };
    std::vector<Item> stack;
    sc.nodes.push_back({});
    stack.push_back({0, 0, n});
    while (!stack.empty()) {
        Item it = stack.back();
        stack.pop_back();
        auto& nd = sc.nodes[it.node];
        V3<R> lo = v3s<R>(std::numeric_limits<R>::infinity()), hi = -lo;
        V3<R> clo = lo, chi = hi;
        for (int k = it.first; k < it.first + it.count; ++k) {
            const auto& G = sc.g[sc.order[k]];
            lo = v3<R>(std::min(lo.x, G.lo.x), std::min(lo.y, G.lo.y), std::min(lo.z, G.lo.z));
            hi = v3<R>(std::max(hi.x, G.hi.x), std::max(hi.y, G.hi.y), std::max(hi.z, G.hi.z));
            clo = v3<R>(std::min(clo.x, G.mu.x), std::min(clo.y, G.mu.y), std::min(clo.z, G.mu.z));
            chi = v3<R>(std::max(chi.x, G.mu.x), std::max(chi.y, G.mu.y), std::max(chi.z, G.mu.z));
        }
        // pad so that rounding never culls a candidate whose peak lies on the box face
        const V3<R> pad = (hi - lo) * R(1e-7) + v3s<R>(R(1e-9));
        nd.lo = lo - pad;
        nd.hi = hi + pad;
        if (it.count <= 4) {
            nd.left = nd.right = -1;
            nd.first = it.first;
            nd.count = it.count;
            continue;
        }
        const V3<R> ce = chi - clo;
        const int axis = (ce.x >= ce.y && ce.x >= ce.z) ? 0 : (ce.y >= ce.z ? 1 : 2);
        const int mid = it.first + it.count / 2;
        std::nth_element(sc.order.begin() + it.first, sc.order.begin() + mid, sc.order.begin() + it.first + it.count,
                         [&](uint32_t a, uint32_t b) {
                             const R ka = axis == 0 ? sc.g[a].mu.x : axis == 1 ? sc.g[a].mu.y : sc.g[a].mu.z;
                             const R kb = axis == 0 ? sc.g[b].mu.x : axis == 1 ? sc.g[b].mu.y : sc.g[b].mu.z;
                             return ka < kb || (ka == kb && a < b);
                         });
        const int l = int(sc.nodes.size());
        sc.nodes.push_back({});
        sc.nodes.push_back({});
        sc.nodes[it.node].left = l;
        sc.nodes[it.node].right = l + 1;
        sc.nodes[it.node].count = 0;
        stack.push_back({l, it.first, mid - it.first});
        stack.push_back({l + 1, mid, it.first + it.count - mid});
    }
}

// Candidate test for Gaussian i (see header): peak t*, inside-cube and distance tests.
// Margin bookkeeping: m = min(m, |a - b| / max(|a|, |b|)) (relative distance to a threshold)
static inline void note_rel(double* m, double a, double b) {
    if (!m) return;
    const double s = std::max(std::fabs(a), std::fabs(b));
    if (s > 0) *m = std::min(*m, std::fabs(a - b) / s);
}
static inline void note_abs(double* m, double d) {
    if (m) *m = std::min(*m, std::fabs(d));
}

template <class R>
static bool candidate(const Gauss<R>& G, const Ray<R>& ray, R& tstar, double* margin = nullptr) {
    const V3<R> dp = ray.o - G.mu;
    // object space of M = R diag(kscl):  x' = diag(1/kscl) R^T x
    const V3<R> rtp = v3<R>(G.rot[0][0] * dp.x + G.rot[1][0] * dp.y + G.rot[2][0] * dp.z,
                            G.rot[0][1] * dp.x + G.rot[1][1] * dp.y + G.rot[2][1] * dp.z,
                            G.rot[0][2] * dp.x + G.rot[1][2] * dp.y + G.rot[2][2] * dp.z);
    const V3<R> rtd = v3<R>(G.rot[0][0] * ray.d.x + G.rot[1][0] * ray.d.y + G.rot[2][0] * ray.d.z,
                            G.rot[0][1] * ray.d.x + G.rot[1][1] * ray.d.y + G.rot[2][1] * ray.d.z,
                            G.rot[0][2] * ray.d.x + G.rot[1][2] * ray.d.y + G.rot[2][2] * ray.d.z);
    const V3<R> op = v3<R>(rtp.x / G.kscl.x, rtp.y / G.kscl.y, rtp.z / G.kscl.z);
    const V3<R> dpr = v3<R>(rtd.x / G.kscl.x, rtd.y / G.kscl.y, rtd.z / G.kscl.z);
    const R dd = dot(dpr, dpr);
    if (!(dd > 0)) return false;
    tstar = -dot(op, dpr) / dd;
    const V3<R> foot = op + dpr * tstar;
    const R fmax_ = std::max(std::fabs(foot.x), std::max(std::fabs(foot.y), std::fabs(foot.z)));
    note_abs(margin, double(fmax_) - 1.0);
    if (fmax_ > 1) return false;
    note_abs(margin, double(tstar));
    const V3<R> gc = cross(safe_normalize(dpr), op);
    note_rel(margin, double(dot(gc, gc) / dd), 9.0);
    return dot(gc, gc) / dd < R(9);  // gaussianParticles.cuh:496-497
}

// All candidates along the ray with t* > 0, sorted by (t*, id).
template <class R>
static void collect(const Scene<R>& sc, const Ray<R>& ray, std::vector<Cand>& out, double* margin = nullptr) {
    out.clear();
    if (sc.nodes.empty()) return;
    const V3<R> inv = v3<R>(R(1) / ray.d.x, R(1) / ray.d.y, R(1) / ray.d.z);
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const auto& nd = sc.nodes[stack[--sp]];
        const V3<R> t0 = (nd.lo - ray.o) * inv, t1 = (nd.hi - ray.o) * inv;
        const R tn = std::fmax(std::fmax(std::fmin(t0.x, t1.x), std::fmin(t0.y, t1.y)), std::fmin(t0.z, t1.z));
        const R tf = std::fmin(std::fmin(std::fmax(t0.x, t1.x), std::fmax(t0.y, t1.y)), std::fmax(t0.z, t1.z));
        if (!(tn <= tf) || tf < 0) continue;
        if (nd.left < 0) {
            for (int k = nd.first; k < nd.first + nd.count; ++k) {
                const uint32_t id = sc.order[k];
                R t;
                if (candidate(sc.g[id], ray, t, margin) && t > 0) out.push_back({double(t), id});
            }
        } else {
            stack[sp++] = nd.left;
            stack[sp++] = nd.right;
        }
    }
    std::sort(out.begin(), out.end(), [](const Cand& a, const Cand& b) { return a.t < b.t || (a.t == b.t && a.id < b.id); });
}

// pathTracer.cuh:46-52 intersectAABB
template <class R>
static void scene_interval(const Scene<R>& sc, const Ray<R>& ray, R& tmin, R& tmax) {
    const R ax = (sc.lo.x - ray.o.x) / ray.d.x, bx = (sc.hi.x - ray.o.x) / ray.d.x;
    const R ay = (sc.lo.y - ray.o.y) / ray.d.y, by = (sc.hi.y - ray.o.y) / ray.d.y;
    const R az = (sc.lo.z - ray.o.z) / ray.d.z, bz = (sc.hi.z - ray.o.z) / ray.d.z;
    const R mx = std::fmax(ax, bx), my = std::fmax(ay, by), mz = std::fmax(az, bz);
    const R nx = std::fmin(ax, bx), ny = std::fmin(ay, by), nz = std::fmin(az, bz);
    tmin = std::fmax(R(0), std::fmax(nx, std::fmax(ny, nz)));
    tmax = std::fmin(mx, std::fmin(my, mz));
}

// One k-buffer chunk: the first `k` candidates with t* in (lo, hi).
static size_t chunk_begin(const std::vector<Cand>& c, double lo) {
    return size_t(std::upper_bound(c.begin(), c.end(), lo, [](double v, const Cand& a) { return v < a.t; }) - c.begin());
}

template <class R>
struct Tracer {
    const Scene<R>& sc;
    const OracleScene& in;
    OracleCounters cnt{};
    std::vector<Cand> cands;
    // gradient sinks (backward only)
    double* dmat = nullptr;
    double* denv = nullptr;
    double* visibility = nullptr;
    const OracleFwdOut* tan_out = nullptr;  // tangents of the outputs (dual-number runs)
    // 3DGRT primary-only tracer: the forward writes the G-buffer only, with the raw
    // weighted sum of shading normals; grt_backward accumulates into these sinks
    bool grt = false;
    double *g_density = nullptr, *g_sh = nullptr, *g_snormal = nullptr;
    // decision margin of the path being traced (see OracleFwdOut::margin)
    mutable double margin = 1.0;
    void note(double a, double b) const { note_rel(&margin, a, b); }
    void note0(double d) const { note_abs(&margin, d); }

    explicit Tracer(const Scene<R>& s) : sc(s), in(*s.in) {}

    const float* sn_of(uint32_t id) const { return in.snormal ? in.snormal + 3 * size_t(id) : nullptr; }

    // pathTracer.cuh:115-128
    bool self_occluded(uint32_t id, R t, const Ray<R>& ray) const {
        const float* sn = sn_of(id);
        if (!sn) return false;
        const R dn = dot(v3<R>(sn[0], sn[1], sn[2]), ray.d);
        if (dn > R(kSelfOccDot)) note(double(t), kMaxSelfOcc);
        if (t > R(kMaxSelfOcc)) return false;
        note(double(dn), kSelfOccDot);
        return dn > R(kSelfOccDot);
    }

    // gaussianParticles.cuh:343-440 (alpha part), also pathTracer.cuh:130-159, 350-374
    R hit_alpha(uint32_t id, const Ray<R>& ray, V3<R>* gro_o = nullptr, V3<R>* grd_o = nullptr, R* gd_o = nullptr) const {
        const Gauss<R>& G = sc.g[id];
        const V3<R> p = ray.o - G.mu;
        const V3<R> pr = v3<R>(G.rot[0][0] * p.x + G.rot[1][0] * p.y + G.rot[2][0] * p.z,
                               G.rot[0][1] * p.x + G.rot[1][1] * p.y + G.rot[2][1] * p.z,
                               G.rot[0][2] * p.x + G.rot[1][2] * p.y + G.rot[2][2] * p.z);
        const V3<R> dr = v3<R>(G.rot[0][0] * ray.d.x + G.rot[1][0] * ray.d.y + G.rot[2][0] * ray.d.z,
                               G.rot[0][1] * ray.d.x + G.rot[1][1] * ray.d.y + G.rot[2][1] * ray.d.z,
                               G.rot[0][2] * ray.d.x + G.rot[1][2] * ray.d.y + G.rot[2][2] * ray.d.z);
        const V3<R> gro = G.is * pr;
        const V3<R> grd = safe_normalize(G.is * dr);
        const V3<R> gc = cross(grd, gro);
        const R gd = dot(gc, gc);
        const R gres = response<R>(int(in.kernel_degree), gd);
        const R a = std::min(R(0.99), gres * G.sigma);
        if (gro_o) { *gro_o = gro; *grd_o = grd; *gd_o = gd; }
        note(double(gres), double(in.min_response));
        note(double(gres * G.sigma), 1.0 / 255.0);
        return (gres > R(in.min_response) && a > R(1.0 / 255.0)) ? a : R(0);
    }

    V3<R> sh_radiance(uint32_t id, V3<R> d) const {
        const float* c = in.sh + size_t(in.sh_stride) * id;
        const size_t cb = size_t(in.sh_stride) * id;
        auto C = [&](int k) {
            return v3<R>(input<R>(c[3 * k], in.t_sh, cb + 3 * k), input<R>(c[3 * k + 1], in.t_sh, cb + 3 * k + 1),
                         input<R>(c[3 * k + 2], in.t_sh, cb + 3 * k + 2));
        };
        // gaussianParticles.cuh:48-87 real SH up to degree 3
        const R C0 = 0.28209479177387814, C1 = 0.4886025119029199;
        const R C2[] = {1.0925484305920792, -1.0925484305920792, 0.31539156525252005, -1.0925484305920792, 0.5462742152960396};
        const R C3[] = {-0.5900435899266435, 2.890611442640554, -0.4570457994644658, 0.3731763325901154,
                        -0.4570457994644658, 1.445305721320277, -0.5900435899266435};
        V3<R> r = C0 * C(0);
        const int deg = in.sph_degree;
        if (deg > 0) {
            const R x = d.x, y = d.y, z = d.z;
            r = r - (C1 * y) * C(1) + (C1 * z) * C(2) - (C1 * x) * C(3);
            if (deg > 1) {
                const R xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
                r = r + (C2[0] * xy) * C(4) + (C2[1] * yz) * C(5) + (C2[2] * (2 * zz - xx - yy)) * C(6) +
                    (C2[3] * xz) * C(7) + (C2[4] * (xx - yy)) * C(8);
                if (deg > 2) {
                    r = r + (C3[0] * y * (3 * xx - yy)) * C(9) + (C3[1] * xy * z) * C(10) +
                        (C3[2] * y * (4 * zz - xx - yy)) * C(11) + (C3[3] * z * (2 * zz - 3 * xx - 3 * yy)) * C(12) +
                        (C3[4] * x * (4 * zz - xx - yy)) * C(13) + (C3[5] * z * (xx - yy)) * C(14) +
                        (C3[6] * x * (xx - 3 * yy)) * C(15);
                }
            }
        }
        r = r + v3s<R>(0.5);
        return v3<R>(std::max(r.x, R(0)), std::max(r.y, R(0)), std::max(r.z, R(0)));
    }

    // pathTracer.cuh:216-311 rayIntersect
    void ray_intersect(const Ray<R> ray, Seg<R>& pl, Pcg32& rng, bool selfOcc, bool primary) {
        R t0, t1;
        scene_interval(sc, ray, t0, t1);
        const R sPdf = pl.scatterPdf, lPdf = pl.lightPdf;
        pl = Seg<R>(ray, std::max(R(0), back(t0)));
        pl.scatterPdf = sPdf;
        pl.lightPdf = lPdf;
        R depth = 0;
        Mat<R> mat;
        V3<R> snAcc = v3s<R>(0);
        const R minT = R(in.min_transmittance);
        cnt.segments++;
        bool gathered = false;
        while (pl.lastHit <= t1 && pl.T > minT) {
            if (!gathered) { collect(sc, ray, cands, &margin); gathered = true; }
            const double lo = double(adv(pl.lastHit)), hi = double(adv(t1));
            size_t b = chunk_begin(cands, lo);
            size_t e = b;
            while (e < cands.size() && int(e - b) < in.kbuf && cands[e].t < hi) ++e;
            if (b == e) break;
            for (size_t k = b; k < e; ++k) {
                note(double(pl.T), double(minT));
                if (!(pl.T > minT)) continue;
                const uint32_t id = cands[k].id;
                const R t = R(cands[k].t);
                cnt.candidates++;
                if (selfOcc && self_occluded(id, t, pl.ray)) {
                    pl.lastHit = std::max(pl.lastHit, t);
                    continue;
                }
                V3<R> gro, grd;
                R gd;
                const R a = hit_alpha(id, pl.ray, &gro, &grd, &gd);
                const bool acc = a > 0;
                if (acc) {
                    const Gauss<R>& G = sc.g[id];
                    R mp[5] = {0, 0, 0, 0, 0};  // no material tensor: the 3DGRT tracer
                    if (in.material) {
                        const float* mp0 = in.material + 5 * size_t(id);
                        for (int q = 0; q < 5; ++q) mp[q] = input<R>(mp0[q], in.t_material, 5 * size_t(id) + q);
                    }
                    const R w = a * pl.T;
                    const R proj = dot(grd, -gro);
                    const V3<R> grds = G.s * grd * proj;
                    const R hitT = std::sqrt(dot(grds, grds));
                    if (primary || in.render_opts == 1) pl.radiance += sh_radiance(id, pl.ray.d) * w;
                    mat.albedo += v3<R>(mp[0], mp[1], mp[2]) * w;
                    mat.rough += R(mp[3]) * w;
                    mat.metal += R(mp[4]) * w;
                    pl.T *= (R(1) - a);
                    depth += hitT * w;
                    pl.hitDist2 += hitT * hitT * w;
                    if (in.enable_normals) {
                        const float* sn = sn_of(id);
                        if (sn)
                            snAcc += w * v3<R>(input<R>(sn[0], in.t_snormal, 3 * size_t(id)),
                                               input<R>(sn[1], in.t_snormal, 3 * size_t(id) + 1),
                                               input<R>(sn[2], in.t_snormal, 3 * size_t(id) + 2));
                        // (R s): the reference multiplies the rotation by the scale vector
                        const V3<R> rs = v3<R>(G.rot[0][0] * G.s.x + G.rot[0][1] * G.s.y + G.rot[0][2] * G.s.z,
                                               G.rot[1][0] * G.s.x + G.rot[1][1] * G.s.y + G.rot[1][2] * G.s.z,
                                               G.rot[2][0] * G.s.x + G.rot[2][1] * G.s.y + G.rot[2][2] * G.s.z);
                        pl.normal += w * safe_normalize((gro + grd * (proj - std::sqrt(R(9) - gd))) * rs);
                    }
                    pl.hit = true;
                    const R op = R(1) - pl.T;
                    pl.distortion = std::max(op * pl.hitDist2 - depth * depth, R(0));
                    if (visibility) visibility[id] = 1;
                    cnt.accepted++;
                    if (primary) cnt.accepted_primary++;
                }
                pl.lastHit = std::max(pl.lastHit, t);
                if (in.enable_hitcounts) pl.hits += acc ? 1 : 0;
            }
        }
        const R op = R(1) - pl.T;
        pl.hitDist = depth;
        pl.snRaw = snAcc;
        pl.valid = false;
        if (pl.hit) note(double(op), 0.5);
        if (pl.hit && op > R(0.5)) {
            // interaction.h:47-68
            const R inv = R(1) / std::max(op, R(1e-12));
            const R dd = depth * inv;
            pl.inter.valid = true;
            pl.inter.pos = pl.ray.o + pl.ray.d * dd;
            pl.inter.sn = snAcc * inv;
            pl.inter.m.albedo = mat.albedo * inv;
            pl.inter.m.rough = mat.rough * inv;
            pl.inter.m.metal = mat.metal * inv;
            pl.inter.g = Mat<R>();
        } else {
            const R u = R(rng.next_1d());
            note(double(u), double(pl.T));
            pl.valid = u < pl.T;
        }
    }

    // pathTracer.cuh:161-214 stochastic shadow ray
    bool shadow_occluded(const Ray<R>& ray, Pcg32& rng, R tmin, R tmax) {
        R t0, t1;
        scene_interval(sc, ray, t0, t1);
        R startT = std::max(tmin, back(t0));
        const R endT = adv(std::min(tmax, t1));
        if (startT >= endT) return false;
        bool gathered = false;
        while (startT < endT) {
            if (!gathered) { collect(sc, ray, cands, &margin); gathered = true; }
            size_t b = chunk_begin(cands, double(adv(startT)));
            size_t e = b;
            while (e < cands.size() && int(e - b) < in.kbuf && cands[e].t < double(endT)) ++e;
            if (b == e) break;
            R batchEnd = startT;
            for (size_t k = b; k < e; ++k) {
                const uint32_t id = cands[k].id;
                const R t = R(cands[k].t);
                if (t > endT) break;
                batchEnd = std::max(batchEnd, t);
                cnt.shadow_candidates++;
                if (self_occluded(id, t, ray)) continue;
                const R a = hit_alpha(id, ray);
                if (a > 0) {
                    const R u = R(rng.next_1d());
                    note(double(u), double(a));
                    if (u < a) return true;
                }
            }
            if (batchEnd <= startT) break;
            startT = batchEnd;
        }
        return false;
    }

    // ---------------------------------------------------------------- environment
    // environment.cuh:123-158 with offset (0,0): rotZ = pi/2, rotX = 0
    V3<R> env_rotate(V3<R> d) const {
        const R sz = std::sin(R(0.5 * kPi)), cz = std::cos(R(0.5 * kPi));
        const V3<R> r = v3<R>(d.x * cz - d.y * sz, d.x * sz + d.y * cz, d.z);
        return r;  // rotX = 0 leaves y,z unchanged
    }
    V3<R> env_inverse_rotate(V3<R> e) const {
        const R sz = std::sin(R(0.5 * kPi)), cz = std::cos(R(0.5 * kPi));
        return v3<R>(e.x * cz + e.y * sz, -e.x * sz + e.y * cz, e.z);
    }
    struct Foot { int x0, x1, y0, y1; R w00, w10, w01, w11; };
    // environment.cuh:49-87
    Foot footprint(R u, R v) const {
        u = u - std::floor(u);
        v = std::min(std::max(v, R(0)), R(1));
        const R x = u * R(in.env_w) - R(0.5), y = v * R(in.env_h) - R(0.5);
        const int x0 = int(std::floor(x)), y0 = int(std::floor(y));
        const R ax = x - R(x0), ay = y - R(y0);
        auto wrap = [](int v, int s) { int w = v % s; return w < 0 ? w + s : w; };
        auto clampi = [](int v, int lo, int hi) { return std::max(lo, std::min(v, hi)); };
        Foot f;
        f.x0 = wrap(x0, in.env_w);
        f.x1 = wrap(x0 + 1, in.env_w);
        f.y0 = clampi(y0, 0, in.env_h - 1);
        f.y1 = clampi(y0 + 1, 0, in.env_h - 1);
        f.w00 = (1 - ax) * (1 - ay);
        f.w10 = ax * (1 - ay);
        f.w01 = (1 - ax) * ay;
        f.w11 = ax * ay;
        return f;
    }
    V3<R> texel(int x, int y) const {
        const size_t o = 4 * (size_t(y) * in.env_w + x);
        const float* t = in.env + o;
        return v3<R>(input<R>(t[0], in.t_env, o), input<R>(t[1], in.t_env, o + 1), input<R>(t[2], in.t_env, o + 2));
    }
    V3<R> bilinear(const Foot& f) const {
        return texel(f.x0, f.y0) * f.w00 + texel(f.x1, f.y0) * f.w10 + texel(f.x0, f.y1) * f.w01 + texel(f.x1, f.y1) * f.w11;
    }
    void equirect_uv(V3<R> dir, R& u, R& v) const {  // environment.cuh:358-366
        const R theta = std::atan2(dir.x, dir.y);
        const R zcl = std::min(R(1), std::max(R(-1), -dir.z));
        const R phi = std::asin(zcl);
        u = (theta + R(kPi)) * (R(0.5) / R(kPi));
        v = R(0.5) + phi / R(kPi);
    }
    V3<R> env_rotated(V3<R> dir) {  // lookup with an already rotated direction
        R u, v;
        equirect_uv(dir, u, v);
        cnt.env_lookups++;
        return bilinear(footprint(u, v));
    }
    V3<R> background(V3<R> rd) {  // environment.cuh:447-454
        if (!sc.has_env()) return v3s<R>(0);
        return env_rotated(env_rotate(rd));
    }
    // environment.cuh:456-473, 101-114, 421-445
    V3<R> background_bwd(V3<R> rd, V3<R> cg, V3<R>* dirGrad) {
        if (!sc.has_env()) return v3s<R>(0);
        const V3<R> dir = env_rotate(rd);
        R u, v;
        equirect_uv(dir, u, v);
        const Foot f = footprint(u, v);
        cnt.env_lookups++;
        const V3<R> bg = bilinear(f);
        if (denv) {
            auto add = [&](int x, int y, R w) {
                double* p = denv + 4 * (size_t(y) * in.env_w + x);
                p[0] += double(w * cg.x);
                p[1] += double(w * cg.y);
                p[2] += double(w * cg.z);
            };
            add(f.x0, f.y0, f.w00);
            add(f.x1, f.y0, f.w10);
            add(f.x0, f.y1, f.w01);
            add(f.x1, f.y1, f.w11);
        }
        cnt.env_atomics++;
        if (dirGrad) {
            const V3<R> un = safe_normalize(rd);
            V3<R> t, b;
            onb(un, t, b);
            const R eps = R(kEnvDirEps);
            const R scl = R(1) / std::sqrt(R(1) + eps * eps);
            const V3<R> tb = env_rotated(env_rotate((un + eps * t) * scl));
            const V3<R> bb = env_rotated(env_rotate((un + eps * b) * scl));
            const R tg = dot((tb - bg) * (R(1) / eps), cg);
            const R bgd = dot((bb - bg) * (R(1) / eps), cg);
            *dirGrad += t * tg + b * bgd;
        }
        return bg;
    }
    // environment.cuh:227-255 alias sampling (4 draws, left-to-right argument order).  With a
    // CDF table the first two draws invert the row marginal and the row's conditional CDF:
    // the chosen index is the first whose inclusive CDF exceeds the draw (a linear scan here).
    V3<R> sample_env(Pcg32& rng, R& pdf) {
        const int nc = in.tab_h * in.tab_w;
        const float* pdfs = in.tab + 2 * nc;
        int cell;
        if (in.tab_kind == 1) {
            const float ur = rng.next_1d();
            int y = 0;
            while (y < in.tab_h - 1 && !(in.tab[nc + size_t(y) * in.tab_w] > ur)) ++y;
            const float uc = rng.next_1d();
            int x = 0;
            while (x < in.tab_w - 1 && !(in.tab[size_t(y) * in.tab_w + x] > uc)) ++x;
            cell = y * in.tab_w + x;
        } else {
            const float* prob = in.tab;
            const float* alias = in.tab + nc;
            const int init = std::min(int(R(rng.next_1d()) * R(nc)), nc - 1);
            const int al = std::max(0, std::min(int(R(alias[init]) + R(0.5)), nc - 1));
            cell = R(rng.next_1d()) < R(prob[init]) ? init : al;
        }
        const int x = cell % in.tab_w, y = cell / in.tab_w;
        const R u1 = R(rng.next_1d());
        const R u2 = R(rng.next_1d());
        R u = (R(x) + u1) / R(in.tab_w);
        R v = (R(y) + u2) / R(in.tab_h);
        u = u - std::floor(u);
        v = std::min(std::max(v, R(0)), R(1));
        const R theta = u * R(2 * kPi) - R(kPi);
        const R phi = (v - R(0.5)) * R(kPi);
        const V3<R> e = v3<R>(std::sin(theta) * std::cos(phi), std::cos(theta) * std::cos(phi), -std::sin(phi));
        pdf = R(pdfs[cell]);
        return env_inverse_rotate(e);
    }
    // environment.cuh:303-328 nearest-cell pdf
    R env_pdf(V3<R> rd) const {
        if (!sc.has_env_light()) return 0;
        const V3<R> dir = env_rotate(rd);
        R u, v;
        equirect_uv(dir, u, v);
        const double xu = double(u) * in.tab_w, yv = double(v) * in.tab_h;
        note0(std::min(std::fabs(xu - std::round(xu)) / in.tab_w, std::fabs(yv - std::round(yv)) / in.tab_h));
        int x = int(u * R(in.tab_w));
        x = x % in.tab_w;
        if (x < 0) x += in.tab_w;
        int y = std::max(0, std::min(int(v * R(in.tab_h)), in.tab_h - 1));
        return R(in.tab[2 * in.tab_h * in.tab_w + y * in.tab_w + x]);
    }

    // ---------------------------------------------------------------- BRDF (material.cuh)
    static R sat(R v) { return std::min(std::max(v, R(0)), R(1)); }
    static V3<R> sat3(V3<R> v) { return v3<R>(sat(v.x), sat(v.y), sat(v.z)); }
    static R clamp_rough(R r) { return std::min(std::max(r, R(kMinRough)), R(1)); }
    static V3<R> nonneg(V3<R> v) { return v3<R>(std::max(v.x, R(0)), std::max(v.y, R(0)), std::max(v.z, R(0))); }
    static V3<R> nonneg_mask(V3<R> v) { return v3<R>(v.x > 0 ? 1 : 0, v.y > 0 ? 1 : 0, v.z > 0 ? 1 : 0); }
    static V3<R> fsn(V3<R> v, V3<R> fb) {  // fast_brdf_safe_normalize
        const R l2 = dot(v, v);
        return l2 > R(kBrdfEps) ? v * (R(1) / std::sqrt(l2)) : fb;
    }
    static V3<R> to_world(V3<R> n, V3<R> l) {
        V3<R> t, b;
        onb(n, t, b);
        return t * l.x + b * l.y + n * l.z;
    }
    static V3<R> sample_diffuse(V3<R> n, R u1, R u2) {
        const R xi = sat(u1), r = std::sqrt(xi), phi = R(2 * kPi) * sat(u2);
        const R z = std::sqrt(std::max(R(0), R(1) - xi));
        return to_world(n, v3<R>(r * std::cos(phi), r * std::sin(phi), z));
    }
    static V3<R> sample_ggx_h(V3<R> n, R rough, R u1, R u2) {
        const R a = rough * rough, a2 = a * a, xi = sat(u1);
        const R den = std::max(R(1) + (a2 - R(1)) * xi, R(kBrdfEps));
        const R ct = std::sqrt(std::max(R(0), (R(1) - xi) / den));
        const R st = std::sqrt(std::max(R(0), R(1) - ct * ct));
        const R phi = R(2 * kPi) * sat(u2);
        return to_world(n, v3<R>(st * std::cos(phi), st * std::sin(phi), ct));
    }

    struct BG {
        V3<R> value = v3s<R>(0), dAlb = v3s<R>(0), dMet = v3s<R>(0), dRough = v3s<R>(0), dDir = v3s<R>(0);
    };

    // material.cuh:448-492 (value) and 514-600 (grads); roughness clamped inside
    BG eval_ls(V3<R> wo, V3<R> n, V3<R> albedo, R metal, R roughness, V3<R> ldir, R& pdf, bool grads) const {
        BG res;
        const V3<R> L = fsn(ldir, n);
        const R NV = sat(dot(n, wo)), NL = sat(dot(n, L));
        pdf = 0;
        if (NV <= 0 || NL <= 0) return res;
        const R rough = clamp_rough(roughness);
        const V3<R> f0 = v3s<R>(0.04) + (albedo - v3s<R>(0.04)) * metal;
        const V3<R> H = fsn(wo + L, n);
        const R NH = sat(dot(n, H)), VH = sat(dot(wo, H));
        const R x = R(1) - sat(VH), x2 = x * x, q = x2 * x2 * x;
        const V3<R> F = f0 + (v3s<R>(1) - f0) * q;
        const R invPi = R(1.0 / kPi);
        const V3<R> diffCol = albedo * (R(1) - metal);
        const R diffScale = NL * invPi;
        const V3<R> diffuse = diffCol * (v3s<R>(1) - F) * diffScale;
        const R al = rough * rough, a2 = al * al;
        const R dDen = std::max(NH * NH * (a2 - R(1)) + R(1), R(kBrdfEps));
        const R D = a2 * invPi / std::max(dDen * dDen, R(kBrdfEps));
        const R k = R(0.5) * rough * rough;
        const R Dv = std::max(NV * (R(1) - k) + k, R(kBrdfEps));
        const R Dl = std::max(NL * (R(1) - k) + k, R(kBrdfEps));
        const R Gv = NV / Dv, Gl = NL / Dl, G = Gv * Gl;
        const R specScale = D * G / std::max(R(4) * NV, R(kBrdfEps));
        const V3<R> spec = F * specScale;
        const R specPdf = VH > 0 ? D * NH / std::max(R(4) * VH, R(kBrdfEps)) : R(0);
        pdf = R(0.5) * (diffScale + specPdf);
        const V3<R> value = diffuse + spec;
        res.value = nonneg(value);
        if (!grads) return res;
        const V3<R> mask = nonneg_mask(value);
        const R dF_dAlb = metal * (R(1) - q);
        if (in.enable_metallic) {
            const V3<R> dF_dMet = (albedo - v3s<R>(0.04)) * (R(1) - q);
            const V3<R> dDiff_dMet = (-albedo * (v3s<R>(1) - F) - diffCol * dF_dMet) * diffScale;
            res.dMet = (dDiff_dMet + dF_dMet * specScale) * mask;
        }
        const V3<R> dDiff_dAlb = ((v3s<R>(1) - F) - albedo * dF_dAlb) * ((R(1) - metal) * diffScale);
        res.dAlb = (dDiff_dAlb + v3s<R>(dF_dAlb * specScale)) * mask;
        const R dD_da2 = invPi * (R(1) / std::max(dDen * dDen, R(kBrdfEps)) -
                                  (R(2) * a2 * NH * NH) / std::max(dDen * dDen * dDen, R(kBrdfEps)));
        const R dD_dr = dD_da2 * (R(4) * rough * rough * rough);
        const R dGv_dk = -NV * (R(1) - NV) / (Dv * Dv);
        const R dGl_dk = -NL * (R(1) - NL) / (Dl * Dl);
        const R dG_dr = (Gl * dGv_dk + Gv * dGl_dk) * rough;
        const R dSpec_dr = (dD_dr * G + D * dG_dr) / std::max(R(4) * NV, R(kBrdfEps));
        const R mr = (roughness > R(kMinRough) && roughness < R(1)) ? R(1) : R(0);
        res.dRough = F * (dSpec_dr * mr) * mask;
        return res;
    }

    // material.cuh:185-219
    R pdf_rough_grad(V3<R> wo, V3<R> n, R roughness, V3<R> ldir) const {
        const V3<R> L = fsn(ldir, n);
        const R NV = sat(dot(n, wo)), NL = sat(dot(n, L));
        if (NV <= 0 || NL <= 0) return 0;
        if (!(roughness > R(kMinRough) && roughness < R(1))) return 0;
        const R rough = clamp_rough(roughness);
        const V3<R> H = fsn(wo + L, n);
        const R NH = sat(dot(n, H)), VH = sat(dot(wo, H));
        if (VH <= 0) return 0;
        const R a2 = rough * rough * rough * rough;
        const R dDen = std::max(NH * NH * (a2 - R(1)) + R(1), R(kBrdfEps));
        const R dD_da2 = R(1.0 / kPi) * (R(1) / std::max(dDen * dDen, R(kBrdfEps)) -
                                         (R(2) * a2 * NH * NH) / std::max(dDen * dDen * dDen, R(kBrdfEps)));
        const R dD_dr = dD_da2 * (R(4) * rough * rough * rough);
        return R(0.5) * dD_dr * NH / std::max(R(4) * VH, R(kBrdfEps));
    }

    // material.cuh:230-286 (grads=false) and 343-425 (grads=true)
    BG sample_throughput(V3<R> wo, V3<R> n, V3<R> albedo, R metal, R roughness, V3<R> rnd, V3<R>& L, R& pdf, bool grads) const {
        const R rough = clamp_rough(roughness);
        BG res;
        L = n;
        pdf = 0;
        V3<R> dDir = v3s<R>(0);
        if (rnd.z < R(0.5)) {
            L = sample_diffuse(n, rnd.x, rnd.y);
            if (!grads && !(dot(n, L) > 0)) return res;
        } else {
            const V3<R> H = sample_ggx_h(n, rough, rnd.x, rnd.y);
            L = H * (R(2) * dot(wo, H)) - wo;
            const bool ok = dot(n, L) > 0;
            if (!grads && !ok) return res;
            if (grads && (roughness > R(kMinRough) && roughness < R(1)) && ok) {
                const R rs = clamp_rough(roughness + R(kRoughGradEps));
                const R step = rs - rough;
                if (step > R(kBrdfEps)) {
                    const V3<R> Hs = sample_ggx_h(n, rs, rnd.x, rnd.y);
                    const V3<R> Ls = Hs * (R(2) * dot(wo, Hs)) - wo;
                    if (dot(n, Ls) > 0) dDir = (Ls - L) / step;
                }
            }
        }
        note0(double(dot(n, L)));
        note0(double(dot(n, wo)));
        if (dot(n, L) <= 0) return res;
        R p;
        const BG bc = eval_ls(wo, n, albedo, metal, roughness, L, p, grads);
        pdf = p;
        if (p <= R(kBrdfEps)) {
            pdf = 0;
            return res;
        }
        const R invPdf = R(1) / std::max(p, R(kBrdfEps));
        const V3<R> value = bc.value * invPdf;
        res.value = nonneg(value);
        res.dDir = dDir;
        if (grads) {
            const V3<R> mask = nonneg_mask(value);
            const R dPdf = pdf_rough_grad(wo, n, roughness, L);
            res.dAlb = bc.dAlb * invPdf * mask;
            res.dRough = (bc.dRough * invPdf - bc.value * (dPdf * invPdf * invPdf)) * mask;
            if (in.enable_metallic) res.dMet = bc.dMet * invPdf * mask;
        }
        return res;
    }

    void shading_frame(const Seg<R>& s, V3<R>& n, V3<R>& wo, V3<R>& albedo, R& metal) const {
        n = fsn(s.inter.sn, v3<R>(0, 0, 1));
        wo = fsn(-s.ray.d, n);
        albedo = sat3(s.inter.m.albedo);
        metal = in.enable_metallic ? sat(s.inter.m.metal) : R(0);
    }

    static R mis(R a, R b) {  // pathTracer.cuh:40-44
        const R a2 = a * a, b2 = b * b;
        return a2 / std::max(a2 + b2, R(1e-6));
    }

    // pathTracer.cuh:461-504
    void sample_nee(Path<R>& p, Pcg32& rng) {
        const Inter<R> it = p.cur.inter;
        if (!sc.has_env_light()) {  // lightSampler.cuh:74-91: empty sample, no draws
            p.emit = Seg<R>();
            return;
        }
        R envPdf;
        const V3<R> wi = sample_env(rng, envPdf);
        const V3<R> Li = background(wi);
        const R lightPdf = envPdf;
        if (lightPdf <= 0 || is_zero(Li)) { p.emit = Seg<R>(); return; }
        note0(double(dot(safe_normalize(it.sn), wi)));
        if (dot(it.sn, wi) <= 0) { p.emit = Seg<R>(); return; }
        V3<R> n, wo, alb;
        R met;
        shading_frame(p.cur, n, wo, alb, met);
        R scatterPdf;
        const BG bc = eval_ls(wo, n, alb, met, clamp_rough(it.m.rough), wi, scatterPdf, false);
        const Ray<R> lr{it.pos + wi * R(kSelfOccOffset), wi};
        if (shadow_occluded(lr, rng, R(0), R(kInfDist))) { p.emit = Seg<R>(); return; }
        p.emit = Seg<R>(lr, 0);
        p.emit.lightPdf = lightPdf;
        p.emit.scatterPdf = scatterPdf;
        p.emit.light = Li;
        p.emit.contribution = p.thr * bc.value * Li / std::max(lightPdf, R(1e-6));
    }

    // environment.cuh:636-715 (renderOpts 0 and 1)
    void accumulate(Path<R>& p) {
        p.cur.contribution = v3s<R>(0);
        const bool first = p.numBounces == 1u;
        const bool mis_on = in.enable_mis != 0;
        const R bMis = (mis_on && first) ? mis(p.cur.scatterPdf, p.cur.lightPdf) : R(1);
        const R lMis = (mis_on && first) ? mis(p.emit.lightPdf, p.emit.scatterPdf) : R(1);
        const V3<R> nee = p.emit.contribution * lMis;
        if (in.render_opts == 1) {
            if (mis_on && first) p.L += nee;
            if (p.numBounces > 0u && p.cur.inter.valid) {
                p.cur.contribution = p.thr * p.cur.radiance;
                if (mis_on && first) p.cur.contribution = p.cur.contribution * bMis;
                p.L += p.cur.contribution;
                p.indirect += p.cur.contribution;
            } else if (p.cur.valid) {
                if (!sc.has_env()) return;
                const V3<R> bg = background(p.cur.ray.d);
                p.thr = p.thr * p.cur.T;
                p.cur.light += bg;
                p.Lnb += bg;
                p.cur.contribution = p.thr * p.cur.light;
                if (mis_on && first) p.cur.contribution = p.cur.contribution * bMis;
                p.L += p.cur.contribution;
                p.direct += p.cur.contribution;
            }
            return;
        }
        if (mis_on && first) {
            p.L += nee;
            p.direct += nee;
        }
        if (p.cur.valid) {
            if (!sc.has_env()) return;
            const V3<R> bg = background(p.cur.ray.d);
            p.thr = p.thr * p.cur.T;
            p.cur.light += bg;
            p.Lnb += bg;
            p.cur.contribution = p.thr * p.cur.light;
            if (mis_on && first) p.cur.contribution = p.cur.contribution * bMis;
            p.cur.radiance += p.cur.contribution;
            p.L += p.cur.contribution;
            if (p.numBounces < 2u) p.direct += p.cur.contribution;
            else p.indirect += p.cur.contribution;
        }
    }

    // environment.cuh:717-808
    void accumulate_bwd(Path<R>& p) {
        p.cur.contribution = v3s<R>(0);
        const bool first = p.numBounces == 1u;
        const bool mis_on = in.enable_mis != 0;
        const R bMis = (mis_on && first) ? mis(p.cur.scatterPdf, p.cur.lightPdf) : R(1);
        const R lMis = (mis_on && first) ? mis(p.emit.lightPdf, p.emit.scatterPdf) : R(1);
        if (in.render_opts == 1) {
            if (mis_on && first) p.L -= p.emit.contribution * lMis;
            if (p.numBounces > 0u && p.cur.inter.valid) {
                p.cur.contribution = p.thr * p.cur.radiance;
                if (mis_on && first) p.cur.contribution = p.cur.contribution * bMis;
                p.L -= p.cur.contribution;
            } else if (p.cur.valid) {
                if (!sc.has_env()) return;
                p.thr = p.thr * p.cur.T;
                V3<R> eg = p.dL * p.thr;
                if (mis_on && first) eg = eg * bMis;
                eg += p.dLnb;
                const V3<R> vis = background_bwd(p.cur.ray.d, eg, &p.cur.dirGrad);
                p.cur.contribution = p.thr * vis;
                if (mis_on && first) p.cur.contribution = p.cur.contribution * bMis;
                p.L -= p.cur.contribution;
            }
            return;
        }
        if (mis_on && first) p.L -= p.emit.contribution * lMis;
        if (p.cur.valid) {
            if (!sc.has_env()) return;
            p.thr = p.thr * p.cur.T;
            V3<R> eg = p.dL * p.thr;
            if (mis_on && first) eg = eg * bMis;
            eg += p.dLnb;
            const V3<R> vis = background_bwd(p.cur.ray.d, eg, &p.cur.dirGrad);
            p.cur.contribution = p.thr * vis;
            if (mis_on && first) p.cur.contribution = p.cur.contribution * bMis;
            p.cur.radiance += p.cur.contribution;
            p.L -= p.cur.contribution;
        }
    }

    // pathTracer.cuh:430-459
    void sample_next(Path<R>& p, Pcg32& rng) {
        const Seg<R> cs = p.cur;
        V3<R> n, wo, alb;
        R met;
        shading_frame(cs, n, wo, alb, met);
        const R ur0 = R(rng.next_1d());
        const R ur1 = R(rng.next_1d());
        const R ur2 = R(rng.next_1d());
        V3<R> L = cs.ray.d;
        R sPdf = 0;
        const BG bt = sample_throughput(wo, n, alb, met, clamp_rough(cs.inter.m.rough), v3<R>(ur0, ur1, ur2), L, sPdf, false);
        const V3<R> nt = p.thr * bt.value;
        if (is_zero(nt)) {
            p.thr = v3s<R>(0);
            p.active = 0;
            return;
        }
        p.thr = nt;
        const Ray<R> nr{cs.inter.pos + safe_normalize(L) * R(kSelfOccOffset), L};
        p.cur = Seg<R>(nr, 0);
        if (in.enable_mis) {
            p.cur.scatterPdf = sPdf;
            p.cur.lightPdf = env_pdf(L);
        }
    }

    // pathTracer.cuh:313-390 material-gradient flush over a replayed segment
    void intersect_bwd(const Ray<R>& ray, R opacity, R maxHit, const Mat<R>& g, bool selfOcc) {
        const R inv = R(1) / std::max(opacity, R(1e-12));
        const V3<R> ga = g.albedo * inv;
        const R gr = g.rough * inv, gm = g.metal * inv;
        R t0, t1;
        scene_interval(sc, ray, t0, t1);
        R startT = std::max(R(0), back(t0));
        const R endT = adv(std::min(maxHit, t1));
        R T = 1;
        const R minT = R(in.min_transmittance);
        bool gathered = false;
        while (startT < endT && T > minT) {
            if (!gathered) { collect(sc, ray, cands, &margin); gathered = true; }
            size_t b = chunk_begin(cands, double(adv(startT)));
            size_t e = b;
            while (e < cands.size() && int(e - b) < in.kbuf && cands[e].t < double(endT)) ++e;
            if (b == e) break;
            for (size_t k = b; k < e; ++k) {
                if (!(T > minT)) continue;
                const uint32_t id = cands[k].id;
                const R t = R(cands[k].t);
                if (selfOcc && self_occluded(id, t, ray)) {
                    startT = std::max(startT, t);
                    continue;
                }
                const R a = hit_alpha(id, ray);
                if (a > 0) {
                    const R w = a * T;
                    if (dmat) {
                        double* q = dmat + 5 * size_t(id);
                        q[0] += double(w * ga.x);
                        q[1] += double(w * ga.y);
                        q[2] += double(w * ga.z);
                        q[3] += double(w * gr);
                        if (in.enable_metallic) q[4] += double(w * gm);
                    }
                    cnt.mat_atomics++;
                    T *= (R(1) - a);
                }
                startT = std::max(startT, t);
            }
        }
    }

    // pathTracer.cuh:392-428
    void pending_bwd(Path<R>& p) {
        Pending<R>& pd = p.pend;
        if (!pd.valid) return;
        const R dR = dot(p.cur.dirGrad, pd.dDir);
        const Pending<R> c = pd;
        p.pend = Pending<R>();
        if (dR == 0) return;
        Mat<R> g;
        g.rough = dR;
        intersect_bwd(c.ray, c.opacity, c.maxHit, g, c.numBounces > 1u);
    }

    // pathTracer.cuh:506-548
    void nee_grad_bwd(Path<R>& p, const Seg<R>& cs) {
        if (p.numBounces != 1u) return;
        if (p.emit.lightPdf <= 0) return;
        V3<R> n, wo, alb;
        R met;
        shading_frame(cs, n, wo, alb, met);
        R sPdf;
        const BG bc = eval_ls(wo, n, alb, met, cs.inter.m.rough, p.emit.ray.d, sPdf, true);
        const R lMis = mis(p.emit.lightPdf, p.emit.scatterPdf);
        const R scale = lMis / std::max(p.emit.lightPdf, R(1e-6));
        const V3<R> bg = background(p.emit.ray.d);
        const V3<R> dLdB = p.dL * p.thr * bg * scale;
        p.cur.inter.g.albedo += dLdB * bc.dAlb;
    }

    // pathTracer.cuh:551-634
    void sample_next_bwd(Path<R>& p, Pcg32& rng) {
        const Seg<R> cs = p.cur;
        V3<R> n, wo, alb;
        R met;
        shading_frame(cs, n, wo, alb, met);
        const R ur0 = R(rng.next_1d());
        const R ur1 = R(rng.next_1d());
        const R ur2 = R(rng.next_1d());
        V3<R> L = cs.ray.d;
        R sPdf = 0;
        const BG bt = sample_throughput(wo, n, alb, met, cs.inter.m.rough, v3<R>(ur0, ur1, ur2), L, sPdf, true);
        const V3<R> nt = p.thr * bt.value;
        if (is_zero(nt)) {
            p.thr = v3s<R>(0);
            p.active = 0;
            return;
        }
        const bool mis_on = in.enable_mis != 0;
        R lPdf = 0;
        bool firstNee = false;
        if (mis_on) {
            lPdf = env_pdf(L);
            firstNee = p.numBounces == 1u && vmax(nt) >= R(1e-4);
            if (firstNee) nee_grad_bwd(p, cs);
        }
        V3<R> dep = p.L;
        if (mis_on && firstNee) dep -= p.emit.contribution * mis(p.emit.lightPdf, p.emit.scatterPdf);
        const V3<R> num = p.dL * dep;
        const V3<R> dLdB = v3<R>(bt.value.x > R(kBrdfEps) ? num.x / bt.value.x : R(0),
                                 bt.value.y > R(kBrdfEps) ? num.y / bt.value.y : R(0),
                                 bt.value.z > R(kBrdfEps) ? num.z / bt.value.z : R(0));
        p.cur.inter.g.albedo += dLdB * bt.dAlb;
        p.cur.inter.g.rough += dot(dLdB, bt.dRough);
        if (in.enable_metallic) p.cur.inter.g.metal += dot(dLdB, bt.dMet);
        // PendingRayDirectionGrad::set
        p.pend.valid = dot(bt.dDir, bt.dDir) > 0;
        p.pend.ray = cs.ray;
        p.pend.opacity = R(1) - cs.T;
        p.pend.maxHit = cs.lastHit;
        p.pend.dDir = bt.dDir;
        p.pend.numBounces = p.numBounces;
        intersect_bwd(cs.ray, R(1) - cs.T, cs.lastHit, p.cur.inter.g, p.numBounces > 1u);
        p.thr = nt;
        const Ray<R> nr{cs.inter.pos + safe_normalize(L) * R(kSelfOccOffset), L};
        p.cur = Seg<R>(nr, 0);
        if (mis_on) {
            p.cur.scatterPdf = sPdf;
            p.cur.lightPdf = lPdf;
        }
    }

    Ray<R> primary_ray(const OracleRays& r, int64_t pix) const {
        const float* o = r.ori + 3 * pix;
        const float* d = r.dir + 3 * pix;
        const float* M = r.ray_to_world;
        Ray<R> ray;
        ray.o = v3<R>(R(M[0]) * o[0] + R(M[1]) * o[1] + R(M[2]) * o[2] + R(M[3]),
                      R(M[4]) * o[0] + R(M[5]) * o[1] + R(M[6]) * o[2] + R(M[7]),
                      R(M[8]) * o[0] + R(M[9]) * o[1] + R(M[10]) * o[2] + R(M[11]));
        ray.d = v3<R>(R(M[0]) * d[0] + R(M[1]) * d[1] + R(M[2]) * d[2], R(M[4]) * d[0] + R(M[5]) * d[1] + R(M[6]) * d[2],
                      R(M[8]) * d[0] + R(M[9]) * d[1] + R(M[10]) * d[2]);
        return ray;
    }

    // Russian roulette after the BRDF sample leaving vertex numBounces >= rr_start (an
    // extension; the reference has none): survive with q = min(1, max throughput), then
    // divide the throughput by q.  The backward replays the same draw.
    bool roulette(Path<R>& p, Pcg32& rng) {
        if (in.rr_start == 0u || p.numBounces < in.rr_start) return true;
        const R q = std::min(R(1), vmax(p.thr));
        const R u = R(rng.next_1d());
        note(double(u), double(q));
        if (!(u < q)) return false;
        p.thr = p.thr * (R(1) / q);
        return true;
    }

    // referenceOptix.cu:27-71
    void forward(const OracleRays& r, int64_t pix, const OracleFwdOut& out) {
        const int64_t W = r.W, H = r.H;
        const uint32_t x = uint32_t(pix % W), y = uint32_t((pix / W) % H),
                       z = uint32_t(pix / (W * H) + r.z_offset);
        Pcg32 rng;
        rng.seed(Pcg32::launch_seed(x, y, z, in.frame), 0);
        Path<R> p;
        p.maxBounces = in.max_bounces;
        cnt.paths++;
        margin = 1.0;
        ray_intersect(primary_ray(r, pix), p.cur, rng, false, true);
        const Seg<R>& c = p.cur;
        auto put = [&](double* OracleFwdOut::*field, int C, int k, R v) {
            if (double* buf = out.*field) buf[pix * C + k] = double(v);
            if (tan_out)
                if (double* tb = tan_out->*field) tb[pix * C + k] = tangent_of(v);
        };
        put(&OracleFwdOut::rad, 3, 0, c.radiance.x); put(&OracleFwdOut::rad, 3, 1, c.radiance.y); put(&OracleFwdOut::rad, 3, 2, c.radiance.z);
        put(&OracleFwdOut::dns, 1, 0, R(1) - c.T);
        put(&OracleFwdOut::hit, 2, 0, c.hitDist); put(&OracleFwdOut::hit, 2, 1, c.lastHit);
        put(&OracleFwdOut::hit2, 1, 0, c.hitDist2);
        put(&OracleFwdOut::distortion, 1, 0, c.distortion);
        if (in.enable_normals) {
            const V3<R> sno = grt ? c.snRaw : c.inter.sn;
            put(&OracleFwdOut::nrm, 3, 0, c.normal.x); put(&OracleFwdOut::nrm, 3, 1, c.normal.y); put(&OracleFwdOut::nrm, 3, 2, c.normal.z);
            put(&OracleFwdOut::snrm, 3, 0, sno.x); put(&OracleFwdOut::snrm, 3, 1, sno.y); put(&OracleFwdOut::snrm, 3, 2, sno.z);
        }
        put(&OracleFwdOut::mat, 5, 0, c.inter.m.albedo.x); put(&OracleFwdOut::mat, 5, 1, c.inter.m.albedo.y); put(&OracleFwdOut::mat, 5, 2, c.inter.m.albedo.z);
        put(&OracleFwdOut::mat, 5, 3, c.inter.m.rough); put(&OracleFwdOut::mat, 5, 4, c.inter.m.metal);
        if (in.enable_hitcounts) put(&OracleFwdOut::hits, 1, 0, c.hits);
        if (grt || !p.cur.inter.valid) {
            if (out.margin) out.margin[pix] = margin;
            return;
        }
        if (in.enable_mis) sample_nee(p, rng);
        for (unsigned depth = 0; depth < in.max_bounces && p.active; ++depth) {
            accumulate(p);
            p.active &= (depth + 1u < p.maxBounces) && p.cur.inter.valid;
            if (!p.active) break;
            p.numBounces = depth + 1u;
            sample_next(p, rng);
            if (!is_zero(p.thr)) note(double(vmax(p.thr)), 1e-4);
            if (vmax(p.thr) < R(1e-4)) break;
            if (!roulette(p, rng)) break;
            ray_intersect(p.cur.ray, p.cur, rng, true, false);
        }
        if (out.margin) out.margin[pix] = margin;
        put(&OracleFwdOut::pbr, 3, 0, p.L.x); put(&OracleFwdOut::pbr, 3, 1, p.L.y); put(&OracleFwdOut::pbr, 3, 2, p.L.z);
        put(&OracleFwdOut::light, 3, 0, p.Lnb.x); put(&OracleFwdOut::light, 3, 1, p.Lnb.y); put(&OracleFwdOut::light, 3, 2, p.Lnb.z);
        put(&OracleFwdOut::comp, 6, 0, p.direct.x); put(&OracleFwdOut::comp, 6, 1, p.direct.y); put(&OracleFwdOut::comp, 6, 2, p.direct.z);
        put(&OracleFwdOut::comp, 6, 3, p.indirect.x); put(&OracleFwdOut::comp, 6, 4, p.indirect.y); put(&OracleFwdOut::comp, 6, 5, p.indirect.z);
    }

    // referenceBwdOptix.cu:27-74
    // Real SH basis function k (0..15) at direction d, as the forward's sh_radiance weighs
    // coefficient k (gaussianParticles.cuh:54-87).
    static R sh_basis(int k, V3<R> d) {
        const R C0 = 0.28209479177387814, C1 = 0.4886025119029199;
        const R C2[] = {1.0925484305920792, -1.0925484305920792, 0.31539156525252005, -1.0925484305920792, 0.5462742152960396};
        const R C3[] = {-0.5900435899266435, 2.890611442640554, -0.4570457994644658, 0.3731763325901154,
                        -0.4570457994644658, 1.445305721320277, -0.5900435899266435};
        const R x = d.x, y = d.y, z = d.z, xx = x * x, yy = y * y, zz = z * z;
        switch (k) {
        case 0: return C0;
        case 1: return -C1 * y;
        case 2: return C1 * z;
        case 3: return -C1 * x;
        case 4: return C2[0] * x * y;
        case 5: return C2[1] * y * z;
        case 6: return C2[2] * (2 * zz - xx - yy);
        case 7: return C2[3] * x * z;
        case 8: return C2[4] * (xx - yy);
        case 9: return C3[0] * y * (3 * xx - yy);
        case 10: return C3[1] * x * y * z;
        case 11: return C3[2] * y * (4 * zz - xx - yy);
        case 12: return C3[3] * z * (2 * zz - 3 * xx - 3 * yy);
        case 13: return C3[4] * x * (4 * zz - xx - yy);
        case 14: return C3[5] * z * (xx - yy);
        default: return C3[6] * x * (xx - 3 * yy);
        }
    }

    // d response / d grayDist times gresGrd, per kernel degree, as the 3DGRT backward writes it
    // (gaussianParticles.cuh:202-244; degree 1 keeps the reference's sqrt factor)
    static R response_grad(int degree, R gd, R gres, R gresGrd) {
        switch (degree) {
        case 8: return R(-0.000685871056241 * 4.0) * gd * gd * gd * gres * gresGrd;
        case 5: return R(-0.0185185185185 * 2.5) * gd * std::sqrt(gd) * gres * gresGrd;
        case 4: return R(-0.0555555555556 * 2.0) * gd * gres * gresGrd;
        case 3: return R(-0.166666666667 * 1.5) * std::sqrt(gd) * gres * gresGrd;
        case 1: return R(-1.5 * 0.5) * std::sqrt(gd) * gres * gresGrd;
        case 0: return gres > 0 ? (R(0.5 * -0.329630334487) / std::sqrt(gd)) * gresGrd : R(0);
        default: return R(-0.5) * gres * gresGrd;
        }
    }

    // threedgrt_tracer referenceBwdOptix.cu:103-184 (raygen) and gaussianParticles.cuh:473-772
    // (processHitBwd): the gradients of the 3DGRT G-buffer outputs with respect to every
    // Gaussian parameter, from the forward's totals.  Candidates are re-collected on
    // (max(0, t0 - eps), min(lastHit, t1) + eps), which in float32 excludes the hits at the
    // last forward distance (inclusive_end keeps them).
    void grt_backward(const OracleRays& r, int64_t pix, const OracleGrtIn& gi) {
        const Ray<R> ray = primary_ray(r, pix);
        cnt.paths++;
        auto v3d = [&](const double* a) { return v3<R>(R(a[3 * pix]), R(a[3 * pix + 1]), R(a[3 * pix + 2])); };
        auto v3f = [&](const float* a) { return a ? v3<R>(R(a[3 * pix]), R(a[3 * pix + 1]), R(a[3 * pix + 2])) : v3s<R>(0); };
        auto f1 = [&](const float* a) { return a ? R(a[pix]) : R(0); };
        const V3<R> radInt = v3d(gi.rad);
        const R trmInt = R(1) - R(gi.dns[pix]);
        const R dInt = R(gi.hit[2 * pix]), maxHit = R(gi.hit[2 * pix + 1]);
        const R m2Int = R(gi.hit2[pix]);
        const R distInt = R(gi.distortion[pix]);
        const V3<R> snInt = v3d(gi.snrm);
        const V3<R> radG = v3f(gi.rad_grad);
        const R trmG = -f1(gi.dns_grad);
        const R dG = f1(gi.hit_grad);
        const R m2G = f1(gi.hit2_grad);
        const R distG = distInt > 0 ? f1(gi.distortion_grad) : R(0);
        const V3<R> snG = v3f(gi.snrm_grad);
        (void)m2Int;
        R t0, t1;
        scene_interval(sc, ray, t0, t1);
        R startT = std::max(R(0), back(t0));
        const R lim = std::min(maxHit, t1);
        const R endT = gi.inclusive_end ? R(std::nextafter(double(lim), 1e300)) : adv(lim);
        R trm = 1, depth = 0, m2 = 0;
        V3<R> rad = v3s<R>(0), sn = v3s<R>(0);
        const R minT = R(in.min_transmittance);
        bool gathered = false;
        while (startT < endT) {
            if (!gathered) { collect(sc, ray, cands, &margin); gathered = true; }
            const size_t b = chunk_begin(cands, double(adv(startT)));
            size_t e = b;
            while (e < cands.size() && int(e - b) < in.kbuf && cands[e].t < double(endT)) ++e;
            if (b == e) break;
            for (size_t k = b; k < e; ++k) {
                const uint32_t id = cands[k].id;
                cnt.candidates++;
                const Gauss<R>& G = sc.g[id];
                const float* pd = in.density + 12 * size_t(id);
                const R q[4] = {R(pd[4]), R(pd[5]), R(pd[6]), R(pd[7])};  // r, x, y, z
                // R^T v: the reference's v * rotation with rows = columns of R
                auto rt = [&](V3<R> v) {
                    return v3<R>(G.rot[0][0] * v.x + G.rot[1][0] * v.y + G.rot[2][0] * v.z,
                                 G.rot[0][1] * v.x + G.rot[1][1] * v.y + G.rot[2][1] * v.z,
                                 G.rot[0][2] * v.x + G.rot[1][2] * v.y + G.rot[2][2] * v.z);
                };
                auto rv = [&](V3<R> v) {
                    return v3<R>(G.rot[0][0] * v.x + G.rot[0][1] * v.y + G.rot[0][2] * v.z,
                                 G.rot[1][0] * v.x + G.rot[1][1] * v.y + G.rot[1][2] * v.z,
                                 G.rot[2][0] * v.x + G.rot[2][1] * v.y + G.rot[2][2] * v.z);
                };
                // d L / d q of L(v * M(q)) given g = dL/d(v * M): M[i][j] = R[j][i]
                auto dquat = [&](V3<R> v, V3<R> g, R out[4]) {
                    const R gv[3][3] = {{g.x * v.x, g.x * v.y, g.x * v.z},
                                        {g.y * v.x, g.y * v.y, g.y * v.z},
                                        {g.z * v.x, g.z * v.y, g.z * v.z}};  // [i][j] = g_i v_j
                    const R w = q[0], x = q[1], y = q[2], z = q[3];
                    out[0] += 2 * z * gv[0][1] - 2 * y * gv[0][2] - 2 * z * gv[1][0] + 2 * x * gv[1][2] +
                              2 * y * gv[2][0] - 2 * x * gv[2][1];
                    out[1] += 2 * y * gv[0][1] + 2 * z * gv[0][2] + 2 * y * gv[1][0] - 4 * x * gv[1][1] +
                              2 * w * gv[1][2] + 2 * z * gv[2][0] - 2 * w * gv[2][1] - 4 * x * gv[2][2];
                    out[2] += -4 * y * gv[0][0] + 2 * x * gv[0][1] - 2 * w * gv[0][2] + 2 * x * gv[1][0] +
                              2 * z * gv[1][2] + 2 * w * gv[2][0] + 2 * z * gv[2][1] - 4 * y * gv[2][2];
                    out[3] += -4 * z * gv[0][0] + 2 * w * gv[0][1] + 2 * x * gv[0][2] - 2 * w * gv[1][0] -
                              4 * z * gv[1][1] + 2 * y * gv[1][2] + 2 * x * gv[2][0] + 2 * y * gv[2][1];
                };
                const V3<R> gposc = ray.o - G.mu;
                const V3<R> gposcr = rt(gposc);
                const V3<R> gro = G.is * gposcr;
                const V3<R> rayDirR = rt(ray.d);
                const V3<R> grdu = G.is * rayDirR;
                const V3<R> grd = safe_normalize(grdu);
                const V3<R> gcrod = cross(grd, gro);
                const R grayDist = dot(gcrod, gcrod);
                const R gres = response<R>(int(in.kernel_degree), grayDist);
                const R galpha = std::min(R(0.99), gres * G.sigma);
                note(double(gres), double(in.min_response));
                note(double(gres * G.sigma), 1.0 / 255.0);
                startT = std::max(startT, R(cands[k].t));
                if (!(gres > R(in.min_response) && galpha > R(1.0 / 255.0))) continue;
                const R proj = dot(grd, -gro);
                const V3<R> grdd = grd * proj;
                const V3<R> grds = G.s * grdd;
                const R gsqdist = dot(grds, grds);
                const R gdist = std::sqrt(gsqdist);
                const R weight = galpha * trm;
                const R nextT = (1 - galpha) * trm;
                const bool last = nextT <= minT;
                depth += weight * gdist;
                const R residD = std::max(last ? R(0) : (dInt - depth) / nextT, R(0));
                m2 += weight * gsqdist;
                const R aHitG = (gdist - residD) * trm * dG;
                const V3<R> grdsG = (gsqdist > 0 ? (grds * weight / gdist) * dG : v3s<R>(0)) +
                                    (R(2) * weight * m2G) * grds + (R(2) * weight * (1 - trmInt) * distG) * grds -
                                    (R(2) * weight * (dInt / std::max(gdist, R(1e-12))) * distG) * grds;
                const V3<R> sclHitG = grdd * grdsG;
                const R grdScaledDot = dot(grdsG * G.s, grd);
                const V3<R> grdHitG = G.s * grdsG * proj - gro * grdScaledDot;
                const V3<R> groHitG = -grd * grdScaledDot;
                const R residTrm = galpha < R(0.999999) ? trmInt / (1 - galpha) : trm;
                const R aDnsG = residTrm * -trmG;
                // SH (gaussianParticles.cuh:94-170): clamped radiance, coefficient gradients
                const float* c = in.sh + size_t(in.sh_stride) * id;
                const int ncoef = (in.sph_degree + 1) * (in.sph_degree + 1);
                V3<R> raw = v3s<R>(0.5);
                for (int kk = 0; kk < ncoef; ++kk)
                    raw += sh_basis(kk, ray.d) * v3<R>(R(c[3 * kk]), R(c[3 * kk + 1]), R(c[3 * kk + 2]));
                const V3<R> col = v3<R>(std::max(raw.x, R(0)), std::max(raw.y, R(0)), std::max(raw.z, R(0)));
                const V3<R> dRGB = v3<R>(raw.x > 0 ? radG.x * weight : R(0), raw.y > 0 ? radG.y * weight : R(0),
                                         raw.z > 0 ? radG.z * weight : R(0));
                if (g_sh)
                    for (int kk = 0; kk < ncoef; ++kk) {
                        const R bk = sh_basis(kk, ray.d);
                        double* dst = g_sh + size_t(in.sh_stride) * id + 3 * kk;
                        dst[0] += double(bk * dRGB.x);
                        dst[1] += double(bk * dRGB.y);
                        dst[2] += double(bk * dRGB.z);
                    }
                rad += weight * col;
                const V3<R> residR = last ? v3s<R>(0) : (radInt - rad) / nextT;
                const V3<R> residRc = v3<R>(std::max(residR.x, R(0)), std::max(residR.y, R(0)), std::max(residR.z, R(0)));
                const float* snp = sn_of(id);
                const V3<R> psn = snp ? v3<R>(R(snp[0]), R(snp[1]), R(snp[2])) : v3s<R>(0);
                sn += weight * psn;
                const V3<R> residSN = last ? v3s<R>(0) : (snInt - sn) / nextT;
                const R aSnG = trm * dot(psn - residSN, snG);
                if (g_snormal) {
                    g_snormal[3 * size_t(id)] += double(weight * snG.x);
                    g_snormal[3 * size_t(id) + 1] += double(weight * snG.y);
                    g_snormal[3 * size_t(id) + 2] += double(weight * snG.z);
                }
                const R common = aHitG + aDnsG + aSnG + trm * dot(col - residRc, radG);
                double* gd = g_density + 12 * size_t(id);
                gd[3] += double(gres * common);
                const R gresG = G.sigma * common;
                const R grayDistG = response_grad(int(in.kernel_degree), grayDist, gres, gresG);
                const V3<R> gcrodG = (R(2) * grayDistG) * gcrod;
                // gcrod = cross(grd, gro)
                const V3<R> grdG = v3<R>(gcrodG.z * gro.y - gcrodG.y * gro.z, gcrodG.x * gro.z - gcrodG.z * gro.x,
                                         gcrodG.y * gro.x - gcrodG.x * gro.y);
                const V3<R> groG = v3<R>(gcrodG.y * grd.z - gcrodG.z * grd.y, gcrodG.z * grd.x - gcrodG.x * grd.z,
                                         gcrodG.x * grd.y - gcrodG.y * grd.x);
                const V3<R> groT = groG + groHitG;
                const V3<R> sclGro = v3<R>(-gposcr.x / (G.s.x * G.s.x), -gposcr.y / (G.s.y * G.s.y),
                                           -gposcr.z / (G.s.z * G.s.z)) * groT;
                const V3<R> gposcrG = G.is * groT;
                const V3<R> gposcG = rv(gposcrG);
                R dq[4] = {0, 0, 0, 0};
                dquat(gposc, gposcrG, dq);
                gd[0] -= double(gposcG.x);
                gd[1] -= double(gposcG.y);
                gd[2] -= double(gposcG.z);
                // grd = safe_normalize(grdu)
                const V3<R> gin = grdG + grdHitG;
                const R l2 = dot(grdu, grdu);
                V3<R> grduG = v3s<R>(0);
                if (l2 > 0) {
                    const R il = R(1) / std::sqrt(l2);
                    grduG = gin * il - grdu * (dot(gin, grdu) * il * il * il);
                }
                const V3<R> sclDir = v3<R>(-rayDirR.x / (G.s.x * G.s.x), -rayDirR.y / (G.s.y * G.s.y),
                                           -rayDirR.z / (G.s.z * G.s.z)) * grduG;
                gd[8] += double(sclHitG.x + sclGro.x + sclDir.x);
                gd[9] += double(sclHitG.y + sclGro.y + sclDir.y);
                gd[10] += double(sclHitG.z + sclGro.z + sclDir.z);
                dquat(ray.d, G.is * grduG, dq);
                for (int qq = 0; qq < 4; ++qq) gd[4 + qq] += double(dq[qq]);
                trm = nextT;
            }
        }
    }

    void backward(const OracleRays& r, int64_t pix, const OracleBwdIn& bi) {
        const int64_t W = r.W, H = r.H;
        const uint32_t x = uint32_t(pix % W), y = uint32_t((pix / W) % H),
                       z = uint32_t(pix / (W * H) + r.z_offset);
        Pcg32 rng;
        rng.seed(Pcg32::launch_seed(x, y, z, in.frame), 0);
        Path<R> p;
        p.maxBounces = in.max_bounces;
        cnt.paths++;
        p.L = v3<R>(R(bi.pbr[3 * pix]), R(bi.pbr[3 * pix + 1]), R(bi.pbr[3 * pix + 2]));
        p.dL = v3<R>(bi.pbr_grad[3 * pix], bi.pbr_grad[3 * pix + 1], bi.pbr_grad[3 * pix + 2]);
        p.Lnb = v3<R>(R(bi.light[3 * pix]), R(bi.light[3 * pix + 1]), R(bi.light[3 * pix + 2]));
        p.dLnb = v3<R>(bi.light_grad[3 * pix], bi.light_grad[3 * pix + 1], bi.light_grad[3 * pix + 2]);
        ray_intersect(primary_ray(r, pix), p.cur, rng, false, true);
        if (p.cur.inter.valid) {
            const float* mg = bi.mat_grad + 5 * pix;
            p.cur.inter.g.albedo += v3<R>(mg[0], mg[1], mg[2]);
            p.cur.inter.g.rough += R(mg[3]);
            if (in.enable_metallic) p.cur.inter.g.metal += R(mg[4]);
        }
        if (!p.cur.inter.valid) return;
        if (in.enable_mis) sample_nee(p, rng);
        for (unsigned depth = 0; depth < in.max_bounces && p.active; ++depth) {
            accumulate_bwd(p);
            pending_bwd(p);
            p.active &= (depth + 1u < p.maxBounces) && p.cur.inter.valid;
            if (!p.active) break;
            p.numBounces = depth + 1u;
            sample_next_bwd(p, rng);
            if (vmax(p.thr) < R(1e-4)) break;
            if (!roulette(p, rng)) break;
            ray_intersect(p.cur.ray, p.cur, rng, true, false);
        }
    }
};

template <class R>
static int run(const OracleScene* in, const OracleRays* rays, const OracleFwdOut* fwd, const OracleBwdIn* bi,
               const OracleBwdOut* bo, int nthreads, OracleCounters* counters, const OracleFwdOut* tangents = nullptr,
               bool grt = false) {
    Scene<R> sc;
    build_scene(sc, in);
    const int64_t total = int64_t(rays->B) * rays->H * rays->W;
    const int64_t stride = std::max<int64_t>(1, rays->path_stride);
    std::vector<int64_t> pix;
    for (int64_t i = rays->path_offset; i < total; i += stride) pix.push_back(i);
    nthreads = std::max(1, nthreads);
    const int n = in->n;
    const size_t envN = (in->env && in->env_h > 0) ? size_t(in->env_h) * in->env_w * 4 : 0;
    std::vector<std::vector<double>> dm(nthreads), de(nthreads);
    std::vector<OracleCounters> cnts(nthreads);
    std::atomic<int64_t> next{0};
    const int64_t grain = 64;
    auto worker = [&](int tid) {
        Tracer<R> tr(sc);
        if (fwd) tr.visibility = fwd->visibility;  // idempotent writes of 1
        tr.tan_out = tangents;
        tr.grt = grt;
        if (bi && bo) {
            if (bo->dmat) { dm[tid].assign(size_t(n) * 5, 0.0); tr.dmat = dm[tid].data(); }
            if (bo->denv && envN) { de[tid].assign(envN, 0.0); tr.denv = de[tid].data(); }
        }
        for (;;) {
            const int64_t b = next.fetch_add(grain);
            if (b >= int64_t(pix.size())) break;
            const int64_t e = std::min<int64_t>(b + grain, int64_t(pix.size()));
            for (int64_t k = b; k < e; ++k) {
                if (bi) tr.backward(*rays, pix[k], *bi);
                else tr.forward(*rays, pix[k], *fwd);
            }
        }
        cnts[tid] = tr.cnt;
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nthreads; ++t) th.emplace_back(worker, t);
    worker(0);
    for (auto& t : th) t.join();
    if (bi && bo) {
        if (bo->dmat) {
            std::fill(bo->dmat, bo->dmat + size_t(n) * 5, 0.0);
            for (int t = 0; t < nthreads; ++t)
                if (!dm[t].empty())
                    for (size_t i = 0; i < dm[t].size(); ++i) bo->dmat[i] += dm[t][i];
        }
        if (bo->denv && envN) {
            std::fill(bo->denv, bo->denv + envN, 0.0);
            for (int t = 0; t < nthreads; ++t)
                if (!de[t].empty())
                    for (size_t i = 0; i < envN; ++i) bo->denv[i] += de[t][i];
        }
    }
    if (counters) {
        OracleCounters c{};
        for (auto& x : cnts) {
            c.paths += x.paths; c.segments += x.segments; c.candidates += x.candidates; c.accepted += x.accepted;
            c.accepted_primary += x.accepted_primary; c.shadow_candidates += x.shadow_candidates;
            c.env_lookups += x.env_lookups; c.mat_atomics += x.mat_atomics; c.env_atomics += x.env_atomics;
        }
        *counters = c;
    }
    return 0;
}

// 3DGRT backward over the launch: per-thread sinks, summed at the end.
static int run_grt_bwd(const OracleScene* in, const OracleRays* rays, const OracleGrtIn* gi, const OracleGrtOut* go,
                       int nthreads) {
    Scene<double> sc;
    build_scene(sc, in);
    const int64_t total = int64_t(rays->B) * rays->H * rays->W;
    const int64_t stride = std::max<int64_t>(1, rays->path_stride);
    std::vector<int64_t> pix;
    for (int64_t i = rays->path_offset; i < total; i += stride) pix.push_back(i);
    nthreads = std::max(1, nthreads);
    const size_t n = size_t(in->n);
    std::vector<std::vector<double>> gd(nthreads), gs(nthreads), gn(nthreads);
    std::atomic<int64_t> next{0};
    auto worker = [&](int tid) {
        Tracer<double> tr(sc);
        tr.grt = true;
        gd[tid].assign(n * 12, 0.0);
        gs[tid].assign(n * size_t(in->sh_stride), 0.0);
        gn[tid].assign(n * 3, 0.0);
        tr.g_density = gd[tid].data();
        tr.g_sh = gs[tid].data();
        tr.g_snormal = gn[tid].data();
        for (;;) {
            const int64_t b = next.fetch_add(64);
            if (b >= int64_t(pix.size())) break;
            const int64_t e = std::min<int64_t>(b + 64, int64_t(pix.size()));
            for (int64_t k = b; k < e; ++k) tr.grt_backward(*rays, pix[k], *gi);
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nthreads; ++t) th.emplace_back(worker, t);
    worker(0);
    for (auto& t : th) t.join();
    auto sum = [&](std::vector<std::vector<double>>& parts, double* dst, size_t len) {
        if (!dst) return;
        std::fill(dst, dst + len, 0.0);
        for (auto& p : parts)
            for (size_t i = 0; i < len; ++i) dst[i] += p[i];
    };
    sum(gd, go->ddensity, n * 12);
    sum(gs, go->dsh, n * size_t(in->sh_stride));
    sum(gn, go->dsnormal, n * 3);
    return 0;
}

}  // namespace orc

extern "C" {

// precision: 64 = double (the oracle proper), 32 = float (same equations in FP32).
int oracle_trace(const OracleScene* scene, const OracleRays* rays, const OracleFwdOut* out, int precision,
                 int nthreads, OracleCounters* counters) {
    if (!scene || !rays || !out) return 1;
    if (precision == 32) return orc::run<float>(scene, rays, out, nullptr, nullptr, nthreads, counters);
    return orc::run<double>(scene, rays, out, nullptr, nullptr, nthreads, counters);
}

int oracle_trace_bwd(const OracleScene* scene, const OracleRays* rays, const OracleBwdIn* bin, const OracleBwdOut* bout,
                     int precision, int nthreads, OracleCounters* counters) {
    if (!scene || !rays || !bin || !bout) return 1;
    if (precision == 32) return orc::run<float>(scene, rays, nullptr, bin, bout, nthreads, counters);
    return orc::run<double>(scene, rays, nullptr, bin, bout, nthreads, counters);
}

// Forward on dual numbers: the outputs in `out` and their derivatives along the direction
// given by the scene's tangent arrays in `tangents` (same layout; visibility unused).
int oracle_trace_dual(const OracleScene* scene, const OracleRays* rays, const OracleFwdOut* out,
                      const OracleFwdOut* tangents, int nthreads) {
    if (!scene || !rays || !out || !tangents) return 1;
    return orc::run<Dual>(scene, rays, out, nullptr, nullptr, nthreads, nullptr, tangents);
}

// 3DGRT primary-only tracer (threedgrt_tracer/src/kernels/cuda/referenceOptix.cu:105-194):
// the G-buffer outputs of the primary segment, with the shading-normal output the raw sum of
// weight x shading normal; the scene's material may be null.  Written into out's rad, dns,
// hit, hit2, distortion, nrm, snrm, hits, visibility (and margin).
int oracle_grt_trace(const OracleScene* scene, const OracleRays* rays, const OracleFwdOut* out, int nthreads) {
    if (!scene || !rays || !out) return 1;
    return orc::run<double>(scene, rays, out, nullptr, nullptr, nthreads, nullptr, nullptr, true);
}

int oracle_grt_trace_dual(const OracleScene* scene, const OracleRays* rays, const OracleFwdOut* out,
                          const OracleFwdOut* tangents, int nthreads) {
    if (!scene || !rays || !out || !tangents) return 1;
    return orc::run<Dual>(scene, rays, out, nullptr, nullptr, nthreads, nullptr, tangents, true);
}

// 3DGRT backward (threedgrt_tracer/src/kernels/cuda/referenceBwdOptix.cu:103-184).
int oracle_grt_trace_bwd(const OracleScene* scene, const OracleRays* rays, const OracleGrtIn* in,
                         const OracleGrtOut* out, int nthreads) {
    if (!scene || !rays || !in || !out || !out->ddensity) return 1;
    return orc::run_grt_bwd(scene, rays, in, out, nthreads);
}

// Building blocks exposed for the closed-form tests.
void oracle_pcg32(uint64_t seed, uint64_t seq, int n, uint32_t* out) {
    orc::Pcg32 r;
    r.seed(seed, seq);
    for (int i = 0; i < n; ++i) out[i] = r.next_u32();
}
uint64_t oracle_launch_seed(uint32_t x, uint32_t y, uint32_t z, uint32_t f) { return orc::Pcg32::launch_seed(x, y, z, f); }
double oracle_response(int degree, double d2) { return orc::response<double>(degree, d2); }
double oracle_kernel_scale(double density, double min_response, int clamping, double degree) {
    return orc::kernel_scale<double>(density, min_response, clamping != 0, degree);
}

// BRDF helpers: value f*cos (3), pdf, and grads for the structural tests.
void oracle_brdf_eval(const double* wo, const double* n, const double* albedo, double metal, double rough, const double* l,
                      int enable_metallic, double* value, double* pdf, double* dalb, double* drough, double* dmet) {
    OracleScene s{};
    s.enable_metallic = enable_metallic;
    orc::Scene<double> sc;
    sc.in = &s;
    orc::Tracer<double> tr(sc);
    using V = orc::V3<double>;
    double p;
    auto r = tr.eval_ls(V{wo[0], wo[1], wo[2]}, V{n[0], n[1], n[2]}, V{albedo[0], albedo[1], albedo[2]}, metal, rough,
                        V{l[0], l[1], l[2]}, p, true);
    value[0] = r.value.x; value[1] = r.value.y; value[2] = r.value.z;
    *pdf = p;
    dalb[0] = r.dAlb.x; dalb[1] = r.dAlb.y; dalb[2] = r.dAlb.z;
    drough[0] = r.dRough.x; drough[1] = r.dRough.y; drough[2] = r.dRough.z;
    dmet[0] = r.dMet.x; dmet[1] = r.dMet.y; dmet[2] = r.dMet.z;
}

void oracle_brdf_sample(const double* wo, const double* n, const double* albedo, double metal, double rough,
                        const double* rnd, int enable_metallic, double* value, double* dir, double* pdf) {
    OracleScene s{};
    s.enable_metallic = enable_metallic;
    orc::Scene<double> sc;
    sc.in = &s;
    orc::Tracer<double> tr(sc);
    using V = orc::V3<double>;
    V L;
    double p;
    auto r = tr.sample_throughput(V{wo[0], wo[1], wo[2]}, V{n[0], n[1], n[2]}, V{albedo[0], albedo[1], albedo[2]}, metal,
                                  rough, V{rnd[0], rnd[1], rnd[2]}, L, p, false);
    value[0] = r.value.x; value[1] = r.value.y; value[2] = r.value.z;
    dir[0] = L.x; dir[1] = L.y; dir[2] = L.z;
    *pdf = p;
}

// Environment lookups for the closed-form tests: radiance of a world direction,
// nearest-cell pdf, and one alias-table sample from four uniforms.
void oracle_env_lookup(const OracleScene* scene, const double* d, double* rgb) {
    orc::Scene<double> sc;
    sc.in = scene;
    orc::Tracer<double> tr(sc);
    auto c = tr.background(orc::V3<double>{d[0], d[1], d[2]});
    rgb[0] = c.x; rgb[1] = c.y; rgb[2] = c.z;
}
double oracle_env_pdf(const OracleScene* scene, const double* d) {
    orc::Scene<double> sc;
    sc.in = scene;
    orc::Tracer<double> tr(sc);
    return tr.env_pdf(orc::V3<double>{d[0], d[1], d[2]});
}
void oracle_env_sample(const OracleScene* scene, uint64_t seed, int n, double* dirs, double* pdfs) {
    orc::Scene<double> sc;
    sc.in = scene;
    orc::Tracer<double> tr(sc);
    orc::Pcg32 rng;
    rng.seed(seed, 0);
    for (int i = 0; i < n; ++i) {
        double pdf;
        auto d = tr.sample_env(rng, pdf);
        dirs[3 * i] = d.x; dirs[3 * i + 1] = d.y; dirs[3 * i + 2] = d.z;
        pdfs[i] = pdf;
    }
}

// Sorted candidate list (t*, id) of one ray, brute force (use_bvh=0) or through the BVH.
int oracle_candidates(const OracleScene* scene, const double* o, const double* d, int use_bvh, int cap, double* ts,
                      uint32_t* ids) {
    orc::Scene<double> sc;
    orc::build_scene(sc, scene);
    orc::Ray<double> ray{orc::V3<double>{o[0], o[1], o[2]}, orc::V3<double>{d[0], d[1], d[2]}};
    std::vector<orc::Cand> c;
    if (use_bvh) {
        orc::collect(sc, ray, c);
    } else {
        for (int i = 0; i < scene->n; ++i) {
            double t;
            if (orc::candidate(sc.g[i], ray, t) && t > 0) c.push_back({t, uint32_t(i)});
        }
        std::sort(c.begin(), c.end(), [](const orc::Cand& a, const orc::Cand& b) { return a.t < b.t || (a.t == b.t && a.id < b.id); });
    }
    const int m = std::min<int>(cap, int(c.size()));
    for (int i = 0; i < m; ++i) { ts[i] = c[i].t; ids[i] = c[i].id; }
    return int(c.size());
}

void oracle_scene_aabb(const OracleScene* scene, double* lo_hi) {
    orc::Scene<double> sc;
    orc::build_scene(sc, scene);
    lo_hi[0] = sc.lo.x; lo_hi[1] = sc.lo.y; lo_hi[2] = sc.lo.z;
    lo_hi[3] = sc.hi.x; lo_hi[4] = sc.hi.y; lo_hi[5] = sc.hi.z;
}

}  // extern "C"
