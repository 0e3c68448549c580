// Test harness: runs the 3DGRT primary-only tracer's device code (ptir_amd/csrc/grt_path.hpp,
// with the traversal of trace_path.hpp) on the CPU, one lane at a time, over the launch of
// host_scene.hpp: the forward writes f_rad, f_dns, f_hit, f_hit2, f_dist, f_nrm, f_snrm,
// f_hits, f_vis (.f32); then the backward, fed with those outputs and the gradients g_rad,
// g_dns, g_hit (depth), g_hit2, g_dist, g_snrm (.f32), writes d_density, d_sh, d_snormal.
// The forward also fills a hit tape of meta tape_cap hits per pixel; a second backward reads
// it instead of re-tracing and writes t_density, t_sh, t_snormal, and tape_redo.f32 holds the
// share of pixels whose tape said re-trace.
#include "../../ptir_amd/csrc/grt_path.hpp"
#include "host_scene.hpp"

using namespace ptir;

int main(int argc, char** argv) {
    if (argc != 2) {
        fprintf(stderr, "usage: grt_host <dir>\n");
        return 2;
    }
    host::Launch la(argv[1]);
    la.S.max_bounces = 1u;
    const size_t P = la.P;
    const int N = la.N;
    std::vector<float> rad(3 * P), dns(P), hit(2 * P), hit2(P), dist(P), nrm(3 * P), snrm(3 * P), hits(P);
    std::vector<uint32_t> vis(size_t(N), 0u);
    const uint32_t cap = uint32_t(la.m.get("tape_cap", 8));
    std::vector<uint32_t> tape(size_t((grt::tape_bytes(P, cap) + 3) / 4), 0u);
    const grt::FwdOut fo{rad.data(), dns.data(), hit.data(), hit2.data(), dist.data(), nrm.data(), snrm.data(),
                         hits.data(), vis.data(), tape.data(), cap};
    la.S.wave_traversal = int(la.m.get("wave_traversal", 0));
    std::vector<uint32_t> lds(size_t(LdsLayout<16, kStackDeep>::kWords) * kTraceBlock, 0u);
    std::vector<uint32_t> wst(kStackDeep, 0u);  // one-lane waves: one wave stack
    la.for_each_path([&](const PixelId& px, int t) {
        const Lane L{lds.data() + t};
        grt::fwd_path<16, kStackDeep>(la.S, fo, L, la.S.wave_traversal ? wst.data() : nullptr, px.pix);
    });
    host::write_f32("f_rad", rad);
    host::write_f32("f_dns", dns);
    host::write_f32("f_hit", hit);
    host::write_f32("f_hit2", hit2);
    host::write_f32("f_dist", dist);
    host::write_f32("f_nrm", nrm);
    host::write_f32("f_snrm", snrm);
    host::write_f32("f_hits", hits);
    std::vector<float> visf(static_cast<size_t>(N), 0.f);
    for (int i = 0; i < N; ++i) visf[i] = float(vis[i]);
    host::write_f32("f_vis", visf);
    const auto g_rad = host::read_f32("g_rad", P * 3), g_dns = host::read_f32("g_dns", P);
    const auto g_hit = host::read_f32("g_hit", P), g_hit2 = host::read_f32("g_hit2", P);
    const auto g_dist = host::read_f32("g_dist", P), g_snrm = host::read_f32("g_snrm", P * 3);
    std::vector<float> d_density(size_t(N) * 12, 0.f), d_sh(size_t(N) * la.shs, 0.f), d_snormal(size_t(N) * 3, 0.f);
    const grt::BwdIO io{rad.data(),   dns.data(),    hit.data(),    hit2.data(),      dist.data(),
                        snrm.data(),  g_rad.data(),  g_dns.data(),  g_hit.data(),     g_hit2.data(),
                        g_dist.data(), g_snrm.data(), la.dens.data(), la.sn.data(), d_density.data(),
                        d_sh.data(),  d_snormal.data()};
    la.for_each_path([&](const PixelId& px, int t) {
        const Lane L{lds.data() + t};
        grt::bwd_path<16, kStackDeep>(la.S, io, L, la.S.wave_traversal ? wst.data() : nullptr, px.pix);
    });
    host::write_f32("d_density", d_density);
    host::write_f32("d_sh", d_sh);
    host::write_f32("d_snormal", d_snormal);
    std::vector<float> t_density(size_t(N) * 12, 0.f), t_sh(size_t(N) * la.shs, 0.f), t_snormal(size_t(N) * 3, 0.f);
    grt::BwdIO iot = io;
    iot.d_density = t_density.data();
    iot.d_sh = t_sh.data();
    iot.d_snormal = t_snormal.data();
    iot.tape = tape.data();
    iot.tape_cap = cap;
    la.for_each_path([&](const PixelId& px, int t) {
        const Lane L{lds.data() + t};
        grt::bwd_path<16, kStackDeep>(la.S, iot, L, la.S.wave_traversal ? wst.data() : nullptr, px.pix);
    });
    host::write_f32("t_density", t_density);
    host::write_f32("t_sh", t_sh);
    host::write_f32("t_snormal", t_snormal);
    size_t redo = 0;
    for (size_t i = 0; i < P; ++i) redo += (tape[grt::kTapeHeader + i] & grt::kTapeRetrace) ? 1 : 0;
    host::write_f32("tape_redo", std::vector<float>{float(redo) / float(P)});
    printf("%d blocks, traversal stack overflow flag %d\n", la.S.num_blocks, la.error_flag);
    return la.error_flag;
}
