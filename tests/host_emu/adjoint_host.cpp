// Test harness: runs the exact-adjoint device code (ptir_amd/csrc/adjoint_path.hpp, with the
// traversal and path code of trace_path.hpp) on the CPU, one lane at a time, over the launch
// of host_scene.hpp plus the output gradients g_pbr, g_light, g_mat, g_rad, g_dns, g_hit,
// g_snrm, g_comp (.f32), and writes d_density, d_material, d_sh, d_snormal, d_env (.f32).
#include "../../ptir_amd/csrc/adjoint_path.hpp"
#include "host_scene.hpp"

using namespace ptir;

int main(int argc, char** argv) {
    if (argc != 2) {
        fprintf(stderr, "usage: adjoint_host <dir>\n");
        return 2;
    }
    host::Launch la(argv[1]);
    const size_t P = la.P;
    const int N = la.N;
    const auto g_pbr = host::read_f32("g_pbr", P * 3), g_light = host::read_f32("g_light", P * 3);
    const auto g_mat = host::read_f32("g_mat", P * 5), g_rad = host::read_f32("g_rad", P * 3);
    const auto g_dns = host::read_f32("g_dns", P), g_hit = host::read_f32("g_hit", P * 2);
    const auto g_snrm = host::read_f32("g_snrm", P * 3), g_comp = host::read_f32("g_comp", P * 6);
    std::vector<float> d_density(size_t(N) * 12, 0.f), d_material(size_t(N) * 5, 0.f),
        d_sh(size_t(N) * la.shs, 0.f), d_snormal(size_t(N) * 3, 0.f), d_env(size_t(la.eh) * la.ew * 4, 0.f);
    std::vector<float4> scratch(adjoint_scratch_words(P, la.S.max_bounces));
    AdjIO io{g_pbr.data(), g_light.data(), g_mat.data(), g_rad.data(), g_dns.data(), g_hit.data(), g_snrm.data(),
             g_comp.data(), la.dens.data(), d_density.data(), d_material.data(), d_sh.data(), d_snormal.data(),
             d_env.data(), scratch.data()};
    std::vector<uint32_t> lds(size_t(LdsLayout<16, kStackDeep>::kWords) * kTraceBlock, 0u);
    la.for_each_path([&](const PixelId& px, int t) {
        const Lane L{lds.data() + t};
        trace_adjoint_path<16, kStackDeep>(la.S, io, L, px);
    });
    host::write_f32("d_density", d_density);
    host::write_f32("d_material", d_material);
    host::write_f32("d_sh", d_sh);
    host::write_f32("d_snormal", d_snormal);
    host::write_f32("d_env", d_env);
    printf("%d blocks, traversal stack overflow flag %d\n", la.S.num_blocks, la.error_flag);
    return la.error_flag;
}
