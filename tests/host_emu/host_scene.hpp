// Test harness support: the launch of a host run of the tracer's device code, read from the
// files of one directory (written by the tests under tests/):
//   meta.txt      "key value" lines (sizes and launch options)
//   <name>.f32    float32 inputs: density [N,12], material [N,5], sh [N,S], snormal [N,3],
//                 env [h,w,4], table [3,th,tw], ori/dir [B,H,W,3], ray_to_world [12]
// The hierarchy is a median split with one Gaussian per leaf, collapsed into the 4-wide nodes
// the traversal walks by the phases of wide.hpp (as bvh.hip runs them), and the per-leaf
// records are filled as k_records fills them (bvh.hip); the traversal returns the same hits
// for any hierarchy.  The DeviceScene is filled as make_scene (api.cpp) fills it.
#pragma once

#include <cstdio>
#include <cstring>
#include <functional>
#include <vector>

Dim3 threadIdx, blockIdx, blockDim;

namespace host {

using namespace ptir;

inline const char*& data_dir() {
    static const char* d = nullptr;
    return d;
}

inline std::vector<float> read_f32(const char* name, size_t n) {
    std::vector<float> v(n, 0.f);
    char path[1024];
    snprintf(path, sizeof(path), "%s/%s.f32", data_dir(), name);
    FILE* f = fopen(path, "rb");
    if (!f) {
        fprintf(stderr, "missing %s\n", path);
        exit(2);
    }
    const size_t got = fread(v.data(), 4, n, f);
    fclose(f);
    if (got != n) {
        fprintf(stderr, "%s: %zu of %zu floats\n", path, got, n);
        exit(2);
    }
    return v;
}

inline void write_bytes(const char* name, const void* p, size_t bytes) {
    char path[1024];
    snprintf(path, sizeof(path), "%s/%s.f32", data_dir(), name);
    FILE* f = fopen(path, "wb");
    fwrite(p, 1, bytes, f);
    fclose(f);
}
inline void write_f32(const char* name, const std::vector<float>& v) { write_bytes(name, v.data(), 4 * v.size()); }

struct Meta {
    char key[64][64];
    double val[64];
    int n = 0;
    double operator[](const char* k) const {
        for (int i = 0; i < n; ++i)
            if (!strcmp(key[i], k)) return val[i];
        fprintf(stderr, "meta.txt has no %s\n", k);
        exit(2);
    }
    double get(const char* k, double fallback) const {
        for (int i = 0; i < n; ++i)
            if (!strcmp(key[i], k)) return val[i];
        return fallback;
    }
};

struct Box {
    float lo[3], hi[3];
};

// kernelScale (particlePrimitives.cu:27-51), as bvh.hip computes it
inline float kernel_scale(float density, int degree, float min_response, int clamping) {
    const float modulation = clamping ? density : 1.0f;
    const float minR = fminf(min_response / modulation, 0.97f);
    if (degree == 0) return ((1.0f - minR) / 3.0f) / -0.329630334487f;
    const float a = -4.5f / powf(3.0f, float(degree));
    return powf(logf(minR) / a, 1.0f / float(degree));
}

inline void rotation(const float* q, float R[3][3]) {
    const float w = q[0], x = q[1], y = q[2], z = q[3];
    R[0][0] = 1 - 2 * (y * y + z * z); R[0][1] = 2 * (x * y - w * z); R[0][2] = 2 * (x * z + w * y);
    R[1][0] = 2 * (x * y + w * z); R[1][1] = 1 - 2 * (x * x + z * z); R[1][2] = 2 * (y * z - w * x);
    R[2][0] = 2 * (x * z - w * y); R[2][1] = 2 * (y * z + w * x); R[2][2] = 1 - 2 * (x * x + y * y);
}

struct Launch {
    Meta m;
    int N = 0, B = 0, H = 0, W = 0, eh = 0, ew = 0, th = 0, tw = 0, shs = 0;
    size_t P = 0;
    std::vector<float> dens, mat, sh, sn, env, tab, ori, dir, T12, bounds;
    std::vector<float4> nodes, wnodes, grec, srec, prec;
    std::vector<int> leaf_of;
    int error_flag = 0;
    DeviceScene S{};

    explicit Launch(const char* d) {
        data_dir() = d;
        char path[1024];
        snprintf(path, sizeof(path), "%s/meta.txt", d);
        FILE* f = fopen(path, "r");
        if (!f) {
            fprintf(stderr, "missing %s\n", path);
            exit(2);
        }
        while (m.n < 64 && fscanf(f, "%63s %lf", m.key[m.n], &m.val[m.n]) == 2) ++m.n;
        fclose(f);
        N = int(m["N"]), B = int(m["B"]), H = int(m["H"]), W = int(m["W"]);
        eh = int(m["env_h"]), ew = int(m["env_w"]), th = int(m["tab_h"]), tw = int(m["tab_w"]);
        shs = int(m["sh_stride"]);
        P = size_t(B) * H * W;
        dens = read_f32("density", size_t(N) * 12);
        mat = read_f32("material", size_t(N) * 5);
        sh = read_f32("sh", size_t(N) * shs);
        sn = read_f32("snormal", size_t(N) * 3);
        env = read_f32("env", size_t(eh) * ew * 4);
        tab = read_f32("table", size_t(3) * th * tw);
        ori = read_f32("ori", P * 3);
        dir = read_f32("dir", P * 3);
        T12 = read_f32("ray_to_world", 12);
        build();
        fill();
    }

    void build() {
        const int degree = int(m["kernel_degree"]), clamping = int(m["density_clamping"]);
        const float min_response = float(m["min_response"]);
        std::vector<Box> box(N);
        for (int i = 0; i < N; ++i) {  // k_instances
            const float* p = &dens[12 * size_t(i)];
            float R[3][3];
            rotation(p + 4, R);
            const float ks = kernel_scale(p[3], degree, min_response, clamping);
            for (int r = 0; r < 3; ++r) {
                const float ext = fabsf(R[r][0]) * ks * p[8] + fabsf(R[r][1]) * ks * p[9] + fabsf(R[r][2]) * ks * p[10];
                const float pad = 2e-6f * (fabsf(p[r]) + ext) + 1e-7f;
                box[i].lo[r] = p[r] - ext - pad;
                box[i].hi[r] = p[r] + ext + pad;
            }
        }
        std::vector<int> order(N);
        for (int i = 0; i < N; ++i) order[i] = i;
        std::function<int(int, int, Box&)> split = [&](int a, int b, Box& out) -> int {
            if (b - a == 1) {
                leaf_of.push_back(order[a]);
                out = box[order[a]];
                return ~int(leaf_of.size() - 1);
            }
            float clo[3] = {1e30f, 1e30f, 1e30f}, chi[3] = {-1e30f, -1e30f, -1e30f};
            for (int k = a; k < b; ++k)
                for (int r = 0; r < 3; ++r) {
                    clo[r] = fminf(clo[r], dens[12 * size_t(order[k]) + r]);
                    chi[r] = fmaxf(chi[r], dens[12 * size_t(order[k]) + r]);
                }
            int ax = 0;
            for (int r = 1; r < 3; ++r)
                if (chi[r] - clo[r] > chi[ax] - clo[ax]) ax = r;
            const int mid = (a + b) / 2;
            std::nth_element(order.begin() + a, order.begin() + mid, order.begin() + b, [&](int i, int j) {
                const float ki = dens[12 * size_t(i) + ax], kj = dens[12 * size_t(j) + ax];
                return ki < kj || (ki == kj && i < j);
            });
            const int me = int(nodes.size() / 4);
            nodes.resize(nodes.size() + 4);
            Box bl, br;
            const int l = split(a, mid, bl);
            const int r = split(mid, b, br);
            float4* nd = &nodes[4 * size_t(me)];
            nd[0] = make_float4(bl.lo[0], bl.lo[1], bl.lo[2], __int_as_float(l));
            nd[1] = make_float4(bl.hi[0], bl.hi[1], bl.hi[2], __int_as_float(r));
            nd[2] = make_float4(br.lo[0], br.lo[1], br.lo[2], 0.f);
            nd[3] = make_float4(br.hi[0], br.hi[1], br.hi[2], 0.f);
            for (int q = 0; q < 3; ++q) {
                out.lo[q] = fminf(bl.lo[q], br.lo[q]);
                out.hi[q] = fmaxf(bl.hi[q], br.hi[q]);
            }
            return me;
        };
        Box all;
        S.root = N > 1 ? split(0, N, all) : (leaf_of.push_back(0), all = box[0], ~0);
        if (N > 1) {  // the 4-wide hierarchy: the phases of bvh.hip, in order
            const int M = N - 1;
            std::vector<int32_t> parent(M, -1), wmap(M);
            std::vector<uint32_t> flag(M), pre(M);
            for (int i = 0; i < M; ++i)
                for (int c = 0; c < 2; ++c) {
                    const int ch = __float_as_int(nodes[4 * size_t(i) + c].w);
                    if (ch >= 0) parent[ch] = i;
                }
            for (int i = 0; i < M; ++i) flag[i] = wide::is_wide(parent.data(), i);
            uint32_t run = 0;
            for (int i = 0; i < M; ++i) {
                pre[i] = run;
                run += flag[i];
            }
            for (int i = 0; i < M; ++i) wide::number(flag.data(), pre.data(), i, wmap.data());
            wnodes.assign(size_t(wide::kWords) * run, make_float4(NAN, NAN, NAN, NAN));
            for (int i = 0; i < M; ++i) wide::fill(nodes.data(), wmap.data(), i, wnodes.data());
            int need = 0;
            for (int i = 0; i < M; ++i) need = std::max(need, wide::stack_need(parent.data(), wmap.data(), wnodes.data(), i));
            if (need > kStackDeep) {
                fprintf(stderr, "stack need %d exceeds %d\n", need, kStackDeep);
                exit(2);
            }
        }
        bounds = {all.lo[0], all.lo[1], all.lo[2], all.hi[0], all.hi[1], all.hi[2]};
        grec.resize(4 * size_t(N));
        srec.resize(2 * size_t(N));
        prec.resize(N);
        for (int j = 0; j < N; ++j) {  // k_records
            const int i = leaf_of[j];
            const float* p = &dens[12 * size_t(i)];
            float R[3][3];
            rotation(p + 4, R);
            const float is[3] = {1.0f / p[8], 1.0f / p[9], 1.0f / p[10]};
            grec[4 * j + 0] = make_float4(R[0][0] * is[0], R[1][0] * is[0], R[2][0] * is[0], p[0]);
            grec[4 * j + 1] = make_float4(R[0][1] * is[1], R[1][1] * is[1], R[2][1] * is[1], p[1]);
            grec[4 * j + 2] = make_float4(R[0][2] * is[2], R[1][2] * is[2], R[2][2] * is[2], p[2]);
            grec[4 * j + 3] = make_float4(p[3], kernel_scale(p[3], degree, min_response, clamping), __int_as_float(i), 0.f);
            const float* mt = &mat[5 * size_t(i)];
            srec[2 * j] = make_float4(mt[0], mt[1], mt[2], mt[3]);
            srec[2 * j + 1] = make_float4(mt[4], sn[3 * size_t(i)], sn[3 * size_t(i) + 1], sn[3 * size_t(i) + 2]);
            prec[j] = make_float4(R[0][0] * p[8] + R[0][1] * p[9] + R[0][2] * p[10],
                                  R[1][0] * p[8] + R[1][1] * p[9] + R[1][2] * p[10],
                                  R[2][0] * p[8] + R[2][1] * p[9] + R[2][2] * p[10], 1.f);
        }
    }

    void fill() {
        S.grec = grec.data();
        S.srec = srec.data();
        S.prec = prec.data();
        S.nodes = nodes.data();
        S.wnodes = wnodes.data();
        S.wide_walks = int(m.get("wide_walks", 3));
        S.bounds = bounds.data();
        S.sh = sh.data();
        S.env = reinterpret_cast<const float4*>(env.data());
        S.tab = tab.data();
        S.ray_to_world = T12.data();
        S.ray_ori = ori.data();
        S.ray_dir = dir.data();
        S.n = uint32_t(N);
        S.sh_stride = shs;
        S.sph_degree = int(m["sph_degree"]);
        S.env_w = ew;
        S.env_h = eh;
        S.tab_w = tw;
        S.tab_h = th;
        S.tab_kind = int(m["tab_kind"]);
        S.kernel_degree = int(m["kernel_degree"]);
        S.min_response = float(m["min_response"]);
        S.min_transmittance = float(m["min_transmittance"]);
        S.render_opts = unsigned(m["render_opts"]);
        const unsigned mb = unsigned(m["max_bounces"]);
        S.max_bounces = S.render_opts ? (mb < 2u ? mb : 2u) : mb;
        S.frame = unsigned(m["frame"]);
        S.cache_mode = unsigned(m.get("cache_mode", 0));
        S.B = B;
        S.H = H;
        S.W = W;
        S.enable_normals = int(m["enable_normals"]);
        S.enable_hitcounts = 1;
        S.enable_mis = int(m["enable_mis"]);
        S.enable_metallic = int(m["enable_metallic"]);
        S.has_snormal = 1;
        S.wave_traversal = 0;
        S.images = nullptr;
        S.num_images = 0;
        S.rr_start = unsigned(m["rr_start"]);
        S.counters = nullptr;
        S.error_flag = &error_flag;
        tiling(S);
    }

    // Run f(px, lane) for every path of the launch, block by block and lane by lane.
    template <class F>
    void for_each_path(F&& f) {
        for (int b = 0; b < S.num_blocks; ++b) {
            for (int t = 0; t < kTraceBlock; ++t) {
                blockIdx.x = unsigned(b);
                threadIdx.x = unsigned(t);
                const PixelId px = pixel_of(S);
                if (px.ok) f(px, t);
            }
        }
    }
};

}  // namespace host
