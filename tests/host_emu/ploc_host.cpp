// Test harness: the PLOC builder's phases (ptir_amd/csrc/ploc.hpp) run on the CPU, one loop
// per kernel launch (the phases only read what earlier launches wrote, so a loop over the
// indices has the launch's semantics), with an exclusive sum between phases 2 and 4 as the
// builder's device scan does.  The tree is then checked:
//   * the last cluster is node 0 and every one of the n-1 node indices was made once;
//   * every leaf and every node but the root has exactly one parent, and parent links agree
//     with the child links;
//   * a walk from the root reaches every leaf once and no node twice;
//   * every node's two child boxes are the boxes of its children (leaf boxes, or the union of
//     the child's own child boxes);
// and compared with a plain sequential restatement of the same clustering (written here from
// the algorithm, not from ploc.hpp): the same children for every node.
// Input: <dir>/meta.txt ("n <count>"), <dir>/boxes.f32 [n,6] (lo, hi) in leaf order.
// Prints "depth <d> sah <s> iterations <k>" and "ok" or the first failure (exit 1).
#include "../../ptir_amd/csrc/ploc.hpp"

#include <cstdio>
#include <cstdlib>
#include <tuple>
#include <vector>

Dim3 threadIdx, blockIdx, blockDim;

using namespace ptir;

namespace {

struct Box {
    float lo[3], hi[3];
};

float half_area(const Box& b) {
    const float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    return dx * dy + dy * dz + dz * dx;
}

Box join(const Box& a, const Box& b) {
    Box r;
    for (int q = 0; q < 3; ++q) {
        r.lo[q] = fminf(a.lo[q], b.lo[q]);
        r.hi[q] = fmaxf(a.hi[q], b.hi[q]);
    }
    return r;
}

// sequential clustering: nearest by smallest union area within the radius; among equal areas
// the closest position, then the pair whose lower index is even, then the lower index; mutual
// pairs merge in index order, nodes numbered downward from n-2
void sequential(const std::vector<Box>& leaves, int radius, std::vector<int>& left, std::vector<int>& right) {
    const int n = int(leaves.size());
    std::vector<int> c(n);
    std::vector<Box> b = leaves;
    for (int j = 0; j < n; ++j) c[j] = ~j;
    left.assign(n - 1, 0);
    right.assign(n - 1, 0);
    int next = n - 2;
    while (c.size() > 1) {
        const int m = int(c.size());
        std::vector<int> nn(m);
        for (int i = 0; i < m; ++i) {
            int best = -1;
            float bd = 0.f;
            auto key = [&](int j) {  // ties after the area: (|i-j|, min index odd?, min index)
                const int lo = std::min(i, j);
                return std::make_tuple(std::abs(i - j), lo % 2, lo);
            };
            for (int j = std::max(0, i - radius); j <= std::min(m - 1, i + radius); ++j) {
                if (j == i) continue;
                const float d = half_area(join(b[i], b[j]));
                if (best < 0 || d < bd || (d == bd && key(j) < key(best))) {
                    best = j;
                    bd = d;
                }
            }
            nn[i] = best;
        }
        std::vector<int> c2;
        std::vector<Box> b2;
        for (int i = 0; i < m; ++i) {
            const int j = nn[i];
            if (nn[j] == i) {
                if (i < j) {
                    left[next] = c[i];
                    right[next] = c[j];
                    c2.push_back(next--);
                    b2.push_back(join(b[i], b[j]));
                }
            } else {
                c2.push_back(c[i]);
                b2.push_back(b[i]);
            }
        }
        c.swap(c2);
        b.swap(b2);
    }
}

int fail(const char* what, long a, long b) {
    printf("FAIL %s (%ld, %ld)\n", what, a, b);
    return 1;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 2) return 2;
    char path[1024];
    snprintf(path, sizeof(path), "%s/meta.txt", argv[1]);
    FILE* f = fopen(path, "r");
    int n = 0;
    if (!f || fscanf(f, "n %d", &n) != 1 || n < 2) return 2;
    fclose(f);
    std::vector<float> raw(size_t(n) * 6);
    snprintf(path, sizeof(path), "%s/boxes.f32", argv[1]);
    f = fopen(path, "rb");
    if (!f || fread(raw.data(), 4, raw.size(), f) != raw.size()) return 2;
    fclose(f);
    std::vector<Box> leaves(n);
    for (int j = 0; j < n; ++j)
        for (int q = 0; q < 3; ++q) {
            leaves[j].lo[q] = raw[6 * size_t(j) + q];
            leaves[j].hi[q] = raw[6 * size_t(j) + 3 + q];
        }

    // the builder's buffers and launches
    std::vector<int32_t> c[2] = {std::vector<int32_t>(n), std::vector<int32_t>(n)};
    std::vector<float4> b[2] = {std::vector<float4>(2 * size_t(n)), std::vector<float4>(2 * size_t(n))};
    std::vector<int32_t> nn(n, -7), parent(2 * size_t(n) - 1, -7);
    std::vector<unsigned long long> val(n), scan(n);
    std::vector<float4> nodes(4 * size_t(n - 1), make_float4(NAN, NAN, NAN, NAN));
    ploc::State st[2];
    for (int j = 0; j < n; ++j) {  // k_ploc_init
        c[0][j] = ~j;
        b[0][2 * j] = make_float4(leaves[j].lo[0], leaves[j].lo[1], leaves[j].lo[2], 0.f);
        b[0][2 * j + 1] = make_float4(leaves[j].hi[0], leaves[j].hi[1], leaves[j].hi[2], 0.f);
    }
    st[0] = ploc::State{n, n - 2};
    parent[0] = -1;
    int it = 0;
    for (; it < 100000; ++it) {
        const int a = it & 1, z = a ^ 1;
        if (st[a].cnt <= 1) break;
        const ploc::Iter I{c[a].data(), b[a].data(), c[z].data(), b[z].data(), nn.data(), val.data(), scan.data(),
                           &st[a], &st[z], nodes.data(), parent.data(), n};
        for (int i = 0; i < n; ++i) ploc::nearest(I, i);
        for (int i = 0; i < n; ++i) ploc::flags(I, i);
        unsigned long long acc = 0;
        for (int i = 0; i < n; ++i) {
            scan[i] = acc;
            acc += val[i];
        }
        for (int i = 0; i < n; ++i) ploc::compact(I, i);
        if (st[z].cnt >= st[a].cnt) return fail("no progress", st[a].cnt, st[z].cnt);
    }
    const ploc::State fin = st[it & 1];
    if (fin.cnt != 1 || fin.base != -1) return fail("final state", fin.cnt, fin.base);
    if (c[it & 1][0] != 0) return fail("root is not node 0", c[it & 1][0], 0);

    // structure
    std::vector<int> left(n - 1), right(n - 1), seen_node(n - 1, 0), seen_leaf(n, 0);
    for (int k = 0; k < n - 1; ++k) {
        left[k] = __float_as_int(nodes[4 * size_t(k)].w);
        right[k] = __float_as_int(nodes[4 * size_t(k) + 1].w);
        for (int ch : {left[k], right[k]}) {
            if (ch >= n - 1 || ch < -n) return fail("child link out of range", k, ch);
            const size_t pi = ch >= 0 ? size_t(ch) : size_t(n - 1) + size_t(~ch);
            if (parent[pi] != k) return fail("parent link", k, ch);
        }
    }
    std::vector<int> stack = {0};
    int depth = 0;
    std::vector<int> level = {0};
    double sah = 0.0;
    Box rootbox;
    std::vector<Box> nodebox(n - 1);
    // boxes bottom-up (nodes are made after their children: higher index first)
    for (int k = n - 2; k >= 0; --k) {
        Box kids[2];
        for (int s = 0; s < 2; ++s) {
            const int ch = s ? right[k] : left[k];
            kids[s] = ch < 0 ? leaves[~ch] : nodebox[ch];
            if (ch >= 0 && ch <= k) return fail("child made after its parent", k, ch);
            const float4 lo = nodes[4 * size_t(k) + 2 * s], hi = nodes[4 * size_t(k) + 2 * s + 1];
            const float got[6] = {lo.x, lo.y, lo.z, hi.x, hi.y, hi.z};
            for (int q = 0; q < 3; ++q)
                if (got[q] != kids[s].lo[q] || got[3 + q] != kids[s].hi[q]) return fail("child box", k, s);
        }
        nodebox[k] = join(kids[0], kids[1]);
    }
    rootbox = nodebox[0];
    while (!stack.empty()) {
        const int k = stack.back();
        const int d = level.back();
        stack.pop_back();
        level.pop_back();
        if (k < 0) {
            if (seen_leaf[~k]++) return fail("leaf reached twice", ~k, 0);
            depth = std::max(depth, d);
            continue;
        }
        if (seen_node[k]++) return fail("node reached twice", k, 0);
        sah += (half_area(nodebox[k])) / half_area(rootbox);
        stack.push_back(left[k]);
        level.push_back(d + 1);
        stack.push_back(right[k]);
        level.push_back(d + 1);
    }
    for (int j = 0; j < n; ++j)
        if (!seen_leaf[j]) return fail("leaf not reached", j, 0);
    for (int k = 0; k < n - 1; ++k)
        if (!seen_node[k]) return fail("node not reached", k, 0);

    std::vector<int> rl, rr;
    sequential(leaves, ploc::kRadius, rl, rr);
    for (int k = 0; k < n - 1; ++k)
        if (rl[k] != left[k] || rr[k] != right[k]) return fail("differs from the sequential clustering", k, left[k]);
    printf("depth %d sah %.3f iterations %d\nok\n", depth, sah, it);
    return 0;
}
