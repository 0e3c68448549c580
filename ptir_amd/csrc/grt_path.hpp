// 3DGRT primary-only tracer (threedgrt_tracer/) on this build's hierarchy, records and
// k-buffer: the forward is the G-buffer of the primary segment (radiance, opacity, depth and
// last hit distance, second moment, distortion, normal, the raw weighted sum of shading
// normals, hit count, visibility), the backward the reference's per-hit adjoint of it with
// respect to every Gaussian parameter (positions, densities, quaternions, scales, SH
// coefficients, shading normals).
#pragma once

#include "trace_path.hpp"

namespace ptir {
namespace grt {

// threedgrt_tracer/src/kernels/cuda/referenceOptix.cu:105-194 for one pixel.  wstack: the
// wave's node stack for the wave-coherent primary traversal, or null.
template <int K, int SD>
__device__ __forceinline__ void fwd_path(const DeviceScene& S, const FwdOut& out, const Lane& L, uint32_t* wstack,
                                         size_t pix) {
    const Ray ray = primary_ray(S, pix);
    float t0, t1;
    scene_interval(S, ray, t0, t1);
    float lastHit = fmaxf(0.f, t0 - kEpsT);
    float T = 1.f, depth = 0.f, depth2 = 0.f, distortion = 0.f, hits = 0.f;
    V3 rad = v3s(0.f), nrm = v3s(0.f), sn = v3s(0.f);
    const float minT = S.min_transmittance;
    const float dlen = sqrtf(dot(ray.d, ray.d));
    count(S, C_PATHS);
    count(S, C_SEGMENTS);
    const size_t P = size_t(S.B) * S.H * S.W;
    const TapeView tape{out.tape, out.tape_cap, P};
    uint32_t ntape = 0;  // accepted hits (the tape holds the first tape_cap)
    if (out.tape && pix == 0) {
        uint32_t* h = tape.header();
        h[0] = kTapeMagic; h[1] = uint32_t(S.B); h[2] = uint32_t(S.H); h[3] = uint32_t(S.W);
        h[4] = S.frame; h[5] = S.n; h[6] = out.tape_cap; h[7] = 0u;
    }
    // gaussianParticles.cuh:329-410 processHit
    auto proc = [&](float t, int slot) {
        count(S, C_CANDIDATES);
        HitGeom h;
        const float a = hit_alpha(S, slot, ray, h);
        if (a > 0.f) {
            const float w = a * T;
            const float hitT = t * dlen;
            const uint32_t id = orig_id(S, slot);
            rad += sh_radiance(S, id, ray.d) * w;
            T *= (1.f - a);
            depth += hitT * w;
            depth2 += hitT * hitT * w;
            if (S.enable_normals) {
                const float4 m1 = S.srec[2 * slot + 1];
                sn += w * v3(m1.y, m1.z, m1.w);
                const float4 rs = S.prec[slot];
                const float proj = dot(h.grd, -h.gro);
                nrm += w * safe_normalize((h.gro + h.grd * (proj - sqrtf(9.f - h.gd))) * xyz(rs));
            }
            distortion = fmaxf((1.f - T) * depth2 - depth * depth, 0.f);
            if (out.vis[id] == 0u) out.vis[id] = 1u;
            count(S, C_ACCEPTED);
            count(S, C_ACCEPTED_PRIMARY);
            if (out.tape) {
                if (ntape < out.tape_cap) {
                    tape.slot()[size_t(ntape) * P + pix] = uint32_t(slot);
                    tape.dist()[size_t(ntape) * P + pix] = t;
                }
                ++ntape;
            }
        }
        lastHit = fmaxf(lastHit, t);
        if (S.enable_hitcounts) hits += a > 0.f ? 1.f : 0.f;
    };
    if (wstack) {
        // all lanes of the wave arrive here together
        bool done = false;
        for (;;) {
            const bool need = !done && lastHit <= t1 && T > minT;
            if (__ballot(need) == 0) break;
            const int cnt = wave_collect<K, SD>(S, ray, lastHit + kEpsT, t1 + kEpsT, need, L, wstack);
            if (need) {
                if (cnt == 0) done = true;
                for (int k = 0; k < cnt; ++k) {
                    if (!(T > minT)) break;
                    proc(__uint_as_float(L.at(k)), int(L.at(K + k)));
                }
            }
        }
    } else {
        while (lastHit <= t1 && T > minT) {
            const int cnt = collect<K, SD>(S, ray, lastHit + kEpsT, t1 + kEpsT, L);
            if (cnt == 0) break;
            for (int k = 0; k < cnt; ++k) {
                if (!(T > minT)) break;
                proc(__uint_as_float(L.at(k)), int(L.at(K + k)));
            }
        }
    }
    if (out.tape) {
        // The backward processes the candidates in (its start, min(lastHit, t1) + eps): the
        // accepted hits at or past that end leave the tape.  Where the transmittance ended
        // the walk and the end lies past lastHit, candidates the forward never reached may
        // lie below it: the backward re-traces that pixel, as it does one whose hits
        // overflow the tape.
        const float endT = fminf(lastHit, t1) + kEpsT;
        uint32_t n = ntape;
        bool redo = n > out.tape_cap || (!(T > minT) && endT > lastHit);
        if (!redo)
            while (n > 0 && tape.dist()[size_t(n - 1) * P + pix] >= endT) --n;
        tape.count()[pix] = redo ? kTapeRetrace : n;
    }
    float* o = out.rad + 3 * pix;
    o[0] = rad.x; o[1] = rad.y; o[2] = rad.z;
    out.dns[pix] = 1.f - T;
    o = out.hit + 2 * pix;
    o[0] = depth; o[1] = lastHit;
    out.hit2[pix] = depth2;
    out.dist[pix] = distortion;
    o = out.nrm + 3 * pix;
    o[0] = nrm.x; o[1] = nrm.y; o[2] = nrm.z;
    o = out.snrm + 3 * pix;
    o[0] = sn.x; o[1] = sn.y; o[2] = sn.z;
    out.hits[pix] = hits;
}

// d response / d grayDist times gresGrd, per kernel degree (gaussianParticles.cuh:202-244,
// as written there: degree 1 keeps its sqrt factor)
__device__ __forceinline__ float response_grad(int degree, float gd, float gres, float gresGrd) {
    switch (degree) {
    case 8: return (-0.000685871056241f * 4.f) * gd * gd * gd * gres * gresGrd;
    case 5: return (-0.0185185185185f * 2.5f) * gd * sqrtf(gd) * gres * gresGrd;
    case 4: return (-0.0555555555556f * 2.f) * gd * gres * gresGrd;
    case 3: return (-0.166666666667f * 1.5f) * sqrtf(gd) * gres * gresGrd;
    case 1: return (-1.5f * 0.5f) * sqrtf(gd) * gres * gresGrd;
    case 0: return gres > 0.f ? (0.5f * -0.329630334487f * rsqrtf(gd)) * gresGrd : 0.f;
    default: return -0.5f * gres * gresGrd;
    }
}

// real SH basis function k at direction d (the weights sh_radiance gives coefficient k)
__device__ __forceinline__ float sh_basis(int k, V3 d) {
    const float x = d.x, y = d.y, z = d.z, xx = x * x, yy = y * y, zz = z * z;
    switch (k) {
    case 0: return 0.28209479177387814f;
    case 1: return -0.4886025119029199f * y;
    case 2: return 0.4886025119029199f * z;
    case 3: return -0.4886025119029199f * x;
    case 4: return 1.0925484305920792f * x * y;
    case 5: return -1.0925484305920792f * y * z;
    case 6: return 0.31539156525252005f * (2.f * zz - xx - yy);
    case 7: return -1.0925484305920792f * x * z;
    case 8: return 0.5462742152960396f * (xx - yy);
    case 9: return -0.5900435899266435f * y * (3.f * xx - yy);
    case 10: return 2.890611442640554f * x * y * z;
    case 11: return -0.4570457994644658f * y * (4.f * zz - xx - yy);
    case 12: return 0.3731763325901154f * z * (2.f * zz - 3.f * xx - 3.f * yy);
    case 13: return -0.4570457994644658f * x * (4.f * zz - xx - yy);
    case 14: return 1.445305721320277f * z * (xx - yy);
    default: return -0.5900435899266435f * x * (xx - 3.f * yy);
    }
}

// dL/dq of L(v * M(q)) with g = dL/d(v * M), M the reference's rotation (rows = columns of the
// standard rotation matrix of the quaternion (w,x,y,z), taken as given): dL/dM[i][j] = g_i v_j.
__device__ __forceinline__ void quat_grad(V3 v, V3 g, const float q[4], float dq[4]) {
    const float w = q[0], x = q[1], y = q[2], z = q[3];
    const float G00 = g.x * v.x, G01 = g.x * v.y, G02 = g.x * v.z;
    const float G10 = g.y * v.x, G11 = g.y * v.y, G12 = g.y * v.z;
    const float G20 = g.z * v.x, G21 = g.z * v.y, G22 = g.z * v.z;
    dq[0] += 2.f * (z * G01 - y * G02 - z * G10 + x * G12 + y * G20 - x * G21);
    dq[1] += 2.f * (y * G01 + z * G02 + y * G10 + w * G12 + z * G20 - w * G21) - 4.f * x * (G11 + G22);
    dq[2] += 2.f * (x * G01 - w * G02 + x * G10 + z * G12 + w * G20 + z * G21) - 4.f * y * (G00 + G22);
    dq[3] += 2.f * (w * G01 + x * G02 - w * G10 + y * G12 + x * G20 + y * G21) - 4.f * z * (G00 + G11);
}

// Running state of the backward over one pixel's hits, and the forward's totals.
struct BwdPixel {
    V3 radInt, snInt, radG, snG;
    float trmInt, dInt, distInt, trmG, dG, m2G, distG;
    float trm, depth, m2;
    V3 rad, sn;
};

// gaussianParticles.cuh:473-772 processHitBwd for one candidate.  Returns whether the hit
// counts (alpha > 0).  With a record R (4 float4) and its key Rk, the hit's gradients are not
// added here: R holds the per-hit values they derive from (dRGB and weight; the position,
// density, quaternion and scale gradients), Rk the Gaussian, and k_grt_sum folds them.
__device__ __forceinline__ bool hit_bwd(const DeviceScene& S, const BwdIO& io, const Ray& ray, int slot, BwdPixel& b,
                                        float4* R, uint32_t* Rk) {
    HitGeom h;
    const float galpha = hit_alpha(S, slot, ray, h);
    if (!(galpha > 0.f)) return false;
    const uint32_t id = orig_id(S, slot);
    const float* pd = io.density + 12 * size_t(id);
    const V3 s = v3(pd[8], pd[9], pd[10]);
    const V3 is = v3(1.f / s.x, 1.f / s.y, 1.f / s.z);
    const float q[4] = {pd[4], pd[5], pd[6], pd[7]};
    const float sigma = h.sig;
    const float gres = response(S.kernel_degree, h.gd);  // as hit_alpha evaluated it
    const V3 gro = h.gro, grd = h.grd;
    const V3 gposcr = s * gro;  // R^T (o - mu)
    const V3 gposc = ray.o - v3(S.grec[4 * slot].w, S.grec[4 * slot + 1].w, S.grec[4 * slot + 2].w);
    // the direction before normalisation: diag(1/s) R^T d
    const float4* gr = S.grec + 4 * slot;
    const V3 grdu = v3(dot(xyz(gr[0]), ray.d), dot(xyz(gr[1]), ray.d), dot(xyz(gr[2]), ray.d));
    const V3 rayDirR = s * grdu;
    const float proj = dot(grd, -gro);
    const V3 grdd = grd * proj;
    const V3 grds = s * grdd;
    const float gsqdist = dot(grds, grds);
    const float gdist = sqrtf(gsqdist);
    const float weight = galpha * b.trm;
    const float nextT = (1.f - galpha) * b.trm;
    const bool last = nextT <= S.min_transmittance;
    b.depth += weight * gdist;
    const float residD = fmaxf(last ? 0.f : (b.dInt - b.depth) / nextT, 0.f);
    b.m2 += weight * gsqdist;
    const float aHitG = (gdist - residD) * b.trm * b.dG;
    const V3 grdsG = (gsqdist > 0.f ? grds * (weight / gdist * b.dG) : v3s(0.f)) + grds * (2.f * weight * b.m2G) +
                     grds * (2.f * weight * (1.f - b.trmInt) * b.distG) -
                     grds * (2.f * weight * (b.dInt / fmaxf(gdist, 1e-12f)) * b.distG);
    const V3 sclHitG = grdd * grdsG;
    const float grdScaledDot = dot(grdsG * s, grd);
    const V3 grdHitG = s * grdsG * proj - gro * grdScaledDot;
    const V3 groHitG = -grd * grdScaledDot;
    const float residTrm = galpha < 0.999999f ? b.trmInt / (1.f - galpha) : b.trm;
    const float aDnsG = residTrm * -b.trmG;
    // SH: clamped radiance and coefficient gradients (gaussianParticles.cuh:94-170)
    const float* c = S.sh + size_t(S.sh_stride) * id;
    const int ncoef = (S.sph_degree + 1) * (S.sph_degree + 1);
    V3 raw = v3s(0.5f);
#pragma unroll
    for (int k = 0; k < 16; ++k)
        if (k < ncoef) raw += sh_basis(k, ray.d) * v3(c[3 * k], c[3 * k + 1], c[3 * k + 2]);
    const V3 col = v3(fmaxf(raw.x, 0.f), fmaxf(raw.y, 0.f), fmaxf(raw.z, 0.f));
    const V3 dRGB = v3(raw.x > 0.f ? b.radG.x * weight : 0.f, raw.y > 0.f ? b.radG.y * weight : 0.f,
                       raw.z > 0.f ? b.radG.z * weight : 0.f);
    float* dsh = io.d_sh + size_t(S.sh_stride) * id;
    // The lanes of a coherent wave often hold the same Gaussian at the same step: each lane
    // starts its coefficients at its own offset, so one atomic instruction of the wave
    // spreads over different addresses (measured: backward 85.5 -> 75.2 ms on the
    // object-scale rays).
    const int off = int(threadIdx.x & 63u) % ncoef;
    for (int kk = 0; kk < (R ? 0 : ncoef); ++kk) {
        const int k = kk + off < ncoef ? kk + off : kk + off - ncoef;
        const float bk = sh_basis(k, ray.d);
        atomicAdd(dsh + 3 * k, bk * dRGB.x);
        atomicAdd(dsh + 3 * k + 1, bk * dRGB.y);
        atomicAdd(dsh + 3 * k + 2, bk * dRGB.z);
    }
    b.rad += col * weight;
    const V3 residR = last ? v3s(0.f) : (b.radInt - b.rad) * (1.f / nextT);
    const V3 residRc = v3(fmaxf(residR.x, 0.f), fmaxf(residR.y, 0.f), fmaxf(residR.z, 0.f));
    const V3 psn = io.snormal ? v3(io.snormal[3 * size_t(id)], io.snormal[3 * size_t(id) + 1],
                                   io.snormal[3 * size_t(id) + 2])
                              : v3s(0.f);
    b.sn += psn * weight;
    const V3 residSN = last ? v3s(0.f) : (b.snInt - b.sn) * (1.f / nextT);
    const float aSnG = b.trm * dot(psn - residSN, b.snG);
    if (io.d_snormal && !R) {
        float* dn = io.d_snormal + 3 * size_t(id);
        atomicAdd(dn, weight * b.snG.x);
        atomicAdd(dn + 1, weight * b.snG.y);
        atomicAdd(dn + 2, weight * b.snG.z);
    }
    const float common = aHitG + aDnsG + aSnG + b.trm * dot(col - residRc, b.radG);
    float* gd = io.d_density + 12 * size_t(id);
    if (!R) atomicAdd(gd + 3, gres * common);
    const float grayDistG = response_grad(S.kernel_degree, h.gd, gres, sigma * common);
    const V3 gcrod = cross(grd, gro);
    const V3 gcrodG = gcrod * (2.f * grayDistG);
    // gcrod = cross(grd, gro)
    const V3 grdG = v3(gcrodG.z * gro.y - gcrodG.y * gro.z, gcrodG.x * gro.z - gcrodG.z * gro.x,
                       gcrodG.y * gro.x - gcrodG.x * gro.y);
    const V3 groG = v3(gcrodG.y * grd.z - gcrodG.z * grd.y, gcrodG.z * grd.x - gcrodG.x * grd.z,
                       gcrodG.x * grd.y - gcrodG.y * grd.x);
    const V3 groT = groG + groHitG;
    const V3 sclGro = v3(-gposcr.x * is.x * is.x, -gposcr.y * is.y * is.y, -gposcr.z * is.z * is.z) * groT;
    const V3 gposcrG = is * groT;
    // R gposcrG with R the standard rotation matrix of q
    const float w = q[0], x = q[1], y = q[2], z = q[3];
    const V3 r0 = v3(1.f - 2.f * (y * y + z * z), 2.f * (x * y - w * z), 2.f * (x * z + w * y));
    const V3 r1 = v3(2.f * (x * y + w * z), 1.f - 2.f * (x * x + z * z), 2.f * (y * z - w * x));
    const V3 r2 = v3(2.f * (x * z - w * y), 2.f * (y * z + w * x), 1.f - 2.f * (x * x + y * y));
    const V3 gposcG = v3(dot(r0, gposcrG), dot(r1, gposcrG), dot(r2, gposcrG));
    float dq[4] = {0.f, 0.f, 0.f, 0.f};
    quat_grad(gposc, gposcrG, q, dq);
    if (!R) {
        atomicAdd(gd + 0, -gposcG.x);
        atomicAdd(gd + 1, -gposcG.y);
        atomicAdd(gd + 2, -gposcG.z);
    }
    // grd = safe_normalize(grdu)
    const V3 gin = grdG + grdHitG;
    const float l2 = dot(grdu, grdu);
    V3 grduG = v3s(0.f);
    if (l2 > 0.f) {
        const float il = rsqrtf(l2);
        grduG = gin * il - grdu * (dot(gin, grdu) * il * il * il);
    }
    const V3 sclDir = v3(-rayDirR.x * is.x * is.x, -rayDirR.y * is.y * is.y, -rayDirR.z * is.z * is.z) * grduG;
    const V3 gscl = v3(sclHitG.x + sclGro.x + sclDir.x, sclHitG.y + sclGro.y + sclDir.y, sclHitG.z + sclGro.z + sclDir.z);
    quat_grad(ray.d, is * grduG, q, dq);
    if (R) {
        R[0] = make_float4(dRGB.x, dRGB.y, dRGB.z, weight);
        R[1] = make_float4(-gposcG.x, -gposcG.y, -gposcG.z, gres * common);
        R[2] = make_float4(dq[0], dq[1], dq[2], dq[3]);
        R[3] = make_float4(gscl.x, gscl.y, gscl.z, 0.f);
        *Rk = id;
    } else {
        atomicAdd(gd + 8, gscl.x);
        atomicAdd(gd + 9, gscl.y);
        atomicAdd(gd + 10, gscl.z);
        atomicAdd(gd + 4, dq[0]);
        atomicAdd(gd + 5, dq[1]);
        atomicAdd(gd + 6, dq[2]);
        atomicAdd(gd + 7, dq[3]);
    }
    b.trm = nextT;
    return true;
}

// The hits of pixel pix go to its records while they last (nrec counts them), then to atomics.
__device__ __forceinline__ void hit_bwd_pix(const DeviceScene& S, const BwdIO& io, const Ray& ray, int slot, BwdPixel& b,
                                            size_t pix, uint32_t& nrec) {
    float4* R = nullptr;
    uint32_t* Rk = nullptr;
    if (io.rec && nrec < kRecPerPixel) {
        const size_t s = size_t(nrec) * io.paths + pix;
        R = io.rec + 4 * s;
        Rk = io.rec_key + s;
    }
    if (hit_bwd(S, io, ray, slot, b, R, Rk) && R) ++nrec;
}

// threedgrt_tracer/src/kernels/cuda/referenceBwdOptix.cu:103-184 for one pixel: the forward's
// candidates re-collected on (max(0, t0 - eps), min(lastHit, t1) + eps), in order.  wstack:
// the wave's node stack for the wave-coherent traversal (every lane of the wave calls
// together), or null.
template <int K, int SD>
__device__ __forceinline__ void bwd_path(const DeviceScene& S, const BwdIO& io, const Lane& L, uint32_t* wstack,
                                         size_t pix) {
    const Ray ray = primary_ray(S, pix);
    auto v3at = [&](const float* a) { return a ? v3(a[3 * pix], a[3 * pix + 1], a[3 * pix + 2]) : v3s(0.f); };
    BwdPixel b;
    b.radInt = v3at(io.rad);
    b.trmInt = 1.f - io.dns[pix];
    b.dInt = io.hit[2 * pix];
    const float maxHit = io.hit[2 * pix + 1];
    b.distInt = io.dist[pix];
    b.snInt = v3at(io.snrm);
    b.radG = v3at(io.rad_grad);
    b.trmG = io.dns_grad ? -io.dns_grad[pix] : 0.f;
    b.dG = io.hit_grad ? io.hit_grad[pix] : 0.f;
    b.m2G = io.hit2_grad ? io.hit2_grad[pix] : 0.f;
    b.distG = (b.distInt > 0.f && io.dist_grad) ? io.dist_grad[pix] : 0.f;
    b.snG = v3at(io.snrm_grad);
    b.trm = 1.f;
    b.depth = b.m2 = 0.f;
    b.rad = b.sn = v3s(0.f);
    float t0, t1;
    scene_interval(S, ray, t0, t1);
    float startT = fmaxf(0.f, t0 - kEpsT);
    const float endT = fminf(maxHit, t1) + kEpsT;
    count(S, C_PATHS);
    uint32_t nrec = 0;
    // the forward's hit tape, when it is this launch's and holds the pixel's hits
    bool taped = false;
    if (io.tape) {
        const size_t P = size_t(S.B) * S.H * S.W;
        const TapeView tape{const_cast<uint32_t*>(io.tape), io.tape_cap, P};
        const uint32_t* h = tape.header();
        const bool mine = h[0] == kTapeMagic && h[1] == uint32_t(S.B) && h[2] == uint32_t(S.H) &&
                          h[3] == uint32_t(S.W) && h[4] == S.frame && h[5] == S.n && h[6] == io.tape_cap;
        const uint32_t n = mine ? tape.count()[pix] : kTapeRetrace;
        taped = (n & kTapeRetrace) == 0u;
        if (taped)
            for (uint32_t j = 0; j < n; ++j) hit_bwd_pix(S, io, ray, int(tape.slot()[size_t(j) * P + pix]), b, pix, nrec);
    }
    if (wstack) {
        bool done = false;
        for (;;) {
            const bool need = !taped && !done && startT < endT;
            if (__ballot(need) == 0) break;
            const int cnt = wave_collect<K, SD>(S, ray, startT + kEpsT, endT, need, L, wstack);
            if (need) {
                if (cnt == 0) done = true;
                for (int k = 0; k < cnt; ++k) {
                    const float t = __uint_as_float(L.at(k));
                    hit_bwd_pix(S, io, ray, int(L.at(K + k)), b, pix, nrec);
                    startT = fmaxf(startT, t);
                }
            }
        }
    } else if (!taped) {
        while (startT < endT) {
            const int cnt = collect<K, SD>(S, ray, startT + kEpsT, endT, L);
            if (cnt == 0) break;
            for (int k = 0; k < cnt; ++k) {
                const float t = __uint_as_float(L.at(k));
                hit_bwd_pix(S, io, ray, int(L.at(K + k)), b, pix, nrec);
                startT = fmaxf(startT, t);
            }
        }
    }
    if (io.rec) io.rec_cnt[pix] = nrec;
}

// k_grt_sum for Gaussian id: the gradients of its n records (their slots listed at
// bucket[0, n)), added to what the hits beyond the records left by atomics.  Four lanes per
// Gaussian (q = 0..3 takes records q, q + 4, ...), their sums combined in the quad; lane 0
// writes, so no other lane writes the Gaussian's gradients.
__device__ __forceinline__ void sum_records(const DeviceScene& S, const BwdIO& io, uint32_t id, const uint32_t* bucket,
                                            uint32_t n, uint32_t q) {
    const int ncoef = (S.sph_degree + 1) * (S.sph_degree + 1);
    float sh[48];
#pragma unroll
    for (int i = 0; i < 48; ++i) sh[i] = 0.f;
    V3 dn = v3s(0.f), gpos = v3s(0.f), gscl = v3s(0.f);
    float gdens = 0.f, dq[4] = {0.f, 0.f, 0.f, 0.f};
    for (uint32_t i = q; i < n; i += 4) {
        const size_t s = bucket[i];
        const size_t pix = s % io.paths;
        const float4* R = io.rec + 4 * s;
        const float4 r0 = R[0], r1 = R[1], r2 = R[2], r3 = R[3];
        const V3 d = primary_ray(S, pix).d;
        const V3 dRGB = v3(r0.x, r0.y, r0.z);
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            if (k < ncoef) {
                const float bk = sh_basis(k, d);
                sh[3 * k] += bk * dRGB.x;
                sh[3 * k + 1] += bk * dRGB.y;
                sh[3 * k + 2] += bk * dRGB.z;
            }
        }
        if (io.snrm_grad) dn += v3(io.snrm_grad[3 * pix], io.snrm_grad[3 * pix + 1], io.snrm_grad[3 * pix + 2]) * r0.w;
        gpos += v3(r1.x, r1.y, r1.z);
        gdens += r1.w;
        dq[0] += r2.x;
        dq[1] += r2.y;
        dq[2] += r2.z;
        dq[3] += r2.w;
        gscl += v3(r3.x, r3.y, r3.z);
    }
    if (n == 0) return;
    // the quad's four partial sums (lanes q = 0..3 of one Gaussian)
    auto quad = [](float v) {
        v += __shfl_xor(v, 1);
        return v + __shfl_xor(v, 2);
    };
#pragma unroll
    for (int i = 0; i < 48; ++i) sh[i] = quad(sh[i]);
    dn = v3(quad(dn.x), quad(dn.y), quad(dn.z));
    gpos = v3(quad(gpos.x), quad(gpos.y), quad(gpos.z));
    gscl = v3(quad(gscl.x), quad(gscl.y), quad(gscl.z));
    gdens = quad(gdens);
    for (int i = 0; i < 4; ++i) dq[i] = quad(dq[i]);
    if (q != 0) return;
    float* dsh = io.d_sh + size_t(S.sh_stride) * id;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        if (k < ncoef) {
            dsh[3 * k] += sh[3 * k];
            dsh[3 * k + 1] += sh[3 * k + 1];
            dsh[3 * k + 2] += sh[3 * k + 2];
        }
    }
    if (io.d_snormal) {
        float* o = io.d_snormal + 3 * size_t(id);
        o[0] += dn.x;
        o[1] += dn.y;
        o[2] += dn.z;
    }
    float* gd = io.d_density + 12 * size_t(id);
    gd[0] += gpos.x;
    gd[1] += gpos.y;
    gd[2] += gpos.z;
    gd[3] += gdens;
    gd[4] += dq[0];
    gd[5] += dq[1];
    gd[6] += dq[2];
    gd[7] += dq[3];
    gd[8] += gscl.x;
    gd[9] += gscl.y;
    gd[10] += gscl.z;
}

}  // namespace grt
}  // namespace ptir
