// Test harness: the two forward schedules of the tracer's device code, run on the CPU one
// lane at a time over the launch of host_scene.hpp -- the one-lane-per-path body (fwd_path of
// trace_path.hpp) and the stages of the wavefront schedule (wavefront_path.hpp, with the shadow
// rays on the traversal engine of engine.hpp) applied to each path in turn -- and compared: every output, the visibility flags, the radiance-cache records
// and the path tape (header, per-path records, every segment's hit list) must be equal bit for
// bit.  Extra meta.txt keys: tape (0/1), tape_cap (hit entries per path), wave_traversal.
// Prints "mismatches <n>" and exits with 1 when n > 0.
#include "../../ptir_amd/csrc/wavefront_path.hpp"
#include "host_scene.hpp"

using namespace ptir;

namespace {

struct Outs {
    std::vector<float> rad, dns, hit, hit2, dist, nrm, snrm, mat, hits, pbr, light, comp, cache;
    std::vector<uint32_t> vis, tape;
    FwdOut fo{};
    Outs(size_t P, int N, bool want_tape, uint64_t tape_bytes, bool want_cache) {
        rad.assign(3 * P, -1.f); dns.assign(P, -1.f); hit.assign(2 * P, -1.f); hit2.assign(P, -1.f);
        dist.assign(P, -1.f); nrm.assign(3 * P, -1.f); snrm.assign(3 * P, -1.f); mat.assign(5 * P, -1.f);
        hits.assign(P, -1.f); pbr.assign(3 * P, -1.f); light.assign(3 * P, -1.f); comp.assign(6 * P, -1.f);
        vis.assign(size_t(N), 0u);
        if (want_cache) cache.assign(size_t(nrc::kRecFloats) * P, -1.f);
        if (want_tape) tape.assign(tape_bytes / 4, 0u);
        fo.rad = rad.data(); fo.dns = dns.data(); fo.hit = hit.data(); fo.hit2 = hit2.data(); fo.dist = dist.data();
        fo.nrm = nrm.data(); fo.snrm = snrm.data(); fo.mat = mat.data(); fo.hits = hits.data(); fo.pbr = pbr.data();
        fo.light = light.data(); fo.comp = comp.data(); fo.vis = vis.data();
        fo.tape = Tape{want_tape ? tape.data() : nullptr, 0u, 0u};
        fo.cache = want_cache ? cache.data() : nullptr;
    }
};

size_t diff(const char* what, const std::vector<float>& a, const std::vector<float>& b) {
    size_t n = 0;
    for (size_t i = 0; i < a.size(); ++i)
        if (std::memcmp(&a[i], &b[i], 4) != 0) {
            if (n < 3) fprintf(stderr, "%s[%zu]: %.9g vs %.9g\n", what, i, a[i], b[i]);
            ++n;
        }
    return n;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 2) {
        fprintf(stderr, "usage: wavefront_host <dir>\n");
        return 2;
    }
    host::Launch la(argv[1]);
    la.S.wave_traversal = int(la.m.get("wave_traversal", 0));
    const size_t P = la.P;
    const bool want_tape = la.m.get("tape", 0) != 0;
    const uint32_t cap = uint32_t(la.m.get("tape_cap", 64));
    const bool want_cache = la.S.cache_mode != 0u;
    const uint64_t tbytes = tape_bytes(P, la.S.max_bounces, cap);
    Outs A(P, la.N, want_tape, tbytes, want_cache), Bo(P, la.N, want_tape, tbytes, want_cache);
    if (want_tape) {
        A.fo.tape = Tape{A.tape.data(), cap, la.S.max_bounces};
        Bo.fo.tape = Tape{Bo.tape.data(), cap, la.S.max_bounces};
    }
    std::vector<uint32_t> lds(size_t(LdsLayout<16, kStackDeep>::kWords) * kTraceBlock, 0u);
    std::vector<uint32_t> wst(size_t(kTraceBlock / 64) * kStackDeep, 0u);
    // one lane per path, start to end
    la.for_each_path([&](const PixelId& px, int t) {
        const Lane L{lds.data() + t};
        fwd_path<16, kStackDeep>(la.S, A.fo, L, wst.data() + (t / 64) * kStackDeep, px);
    });
    // the wavefront stages, path by path
    // (each stage saves the path's state to a slot and the next loads it back: here slot = path)
    std::vector<float> state(size_t(wf::kWords) * P, 0.f);
    wf::IO io{};
    io.sa = io.sb = state.data();
    io.P = P;
    size_t staged = 0, jobs = 0;
    la.for_each_path([&](const PixelId& px, int t) {
        const Lane L{lds.data() + t};
        const uint32_t pix = uint32_t(px.pix);
        Path p;
        Rng rng;
        wf::Carry c;
        bool go = wf::primary_path<16, kStackDeep>(la.S, Bo.fo, L, wst.data() + (t / 64) * kStackDeep, px, p, rng, c);
        auto hop = [&](int st) {
            wf::save(wf::slot_of(io, st, pix), p, rng, c);
            Path q;
            Rng r;
            wf::Carry cc;
            wf::load(wf::slot_of(io, st, pix), q, r, cc);
            p = q;
            rng = r;
            c = cc;
        };
        if (go && la.S.enable_mis) {
            // k_wf_nee_setup, k_wf_occlusion (the engine on this path's job alone), k_wf_nee_finish
            ++staged;
            hop(0);
            eng::ShadowJob job;
            if (wf::nee_setup(la.S, p, rng, pix, job)) {
                ++jobs;
                uint32_t head = 0u;
                eng::occlusion<16, kStackDeep>(la.S, &job, 1u, &head, L);
                wf::load(wf::slot_of(io, 0, pix), p, rng, c);  // the state as the finish kernel reads it
                wf::nee_apply(p, rng, job);
            }
            go = wf::nee_done(la.S, Bo.fo, pix, p, rng, c);
        }
        for (uint32_t b = 1; go && b < la.S.max_bounces; ++b) {
            ++staged;
            hop(int(b));
            go = wf::bounce_path<16, kStackDeep>(la.S, Bo.fo, L, pix, b, p, rng, c);
        }
    });
    size_t bad = 0;
    bad += diff("rad", A.rad, Bo.rad);
    bad += diff("dns", A.dns, Bo.dns);
    bad += diff("hit", A.hit, Bo.hit);
    bad += diff("hit2", A.hit2, Bo.hit2);
    bad += diff("dist", A.dist, Bo.dist);
    bad += diff("nrm", A.nrm, Bo.nrm);
    bad += diff("snrm", A.snrm, Bo.snrm);
    bad += diff("mat", A.mat, Bo.mat);
    bad += diff("hits", A.hits, Bo.hits);
    bad += diff("pbr", A.pbr, Bo.pbr);
    bad += diff("light", A.light, Bo.light);
    bad += diff("comp", A.comp, Bo.comp);
    if (want_cache) bad += diff("cache", A.cache, Bo.cache);
    for (int i = 0; i < la.N; ++i) bad += A.vis[i] != Bo.vis[i];
    size_t chunks = 0;
    if (want_tape) {
        const Tape ta = A.fo.tape, tb = Bo.fo.tape;
        for (int w = 0; w < 10; ++w) bad += ta.base[w] != tb.base[w];
        for (size_t pix = 0; pix < P; ++pix) {
            const uint32_t* ra = ta.path(pix);
            const uint32_t* rb = tb.path(pix);
            for (uint32_t w = 0; w < ta.rec_words(); ++w)
                if (w != 3u && ra[w] != rb[w]) ++bad;  // word 3: the pool chunk, allocation order
            chunks += ra[3] != kNoChunk;
            for (uint32_t s = 0; s < ta.segs; ++s) {
                const uint32_t* sa = ra + kTapePathWords + kTapeSegWords * s;
                const uint32_t n = sa[3] & 0xFFFFu, off = sa[15];
                for (uint32_t k = 0; k < n; ++k) {
                    const uint2 ea = *ta.entry(ta.hits(pix, P), ta.pool(P), ra[3], off + k);
                    const uint2 eb = *tb.entry(tb.hits(pix, P), tb.pool(P), rb[3], off + k);
                    bad += ea.x != eb.x || ea.y != eb.y;
                }
            }
        }
    }
    printf("paths %zu, stage runs %zu, shadow jobs %zu, pool chunks %zu, stack overflow %d, mismatches %zu\n", P, staged,
           jobs, chunks, la.error_flag, bad);
    return bad || la.error_flag ? 1 : 0;
}
