#include "group.h"

#include <algorithm>
#include <numeric>
#include <stdexcept>

namespace qsa {

std::vector<std::uint64_t> GroupSpec::byte_lut() const {
    const int N = n_sites, B = bpw();
    std::vector<std::uint64_t> lut(static_cast<std::size_t>(G) * B * 256, 0ULL);
    for (int g = 0; g < G; ++g) {
        const int* p = perms.data() + static_cast<std::size_t>(g) * N;
        int inv[64];
        for (int i = 0; i < N; ++i) inv[p[i]] = i;   // source bit j lands on output bit inv[j]
        for (int b = 0; b < B; ++b)
            for (int v = 0; v < 256; ++v) {
                std::uint64_t o = 0;
                for (int t = 0; t < 8 && 8 * b + t < N; ++t)
                    if ((v >> t) & 1) o |= 1ULL << inv[8 * b + t];
                lut[(static_cast<std::size_t>(g) * B + b) * 256 + v] = o;
            }
    }
    return lut;
}

bool GroupSpec::real_characters() const {
    for (const auto& c : chars)
        if (c.imag() != 0.0) return false;
    return true;
}

namespace {

// The block system with the largest blocks (2..8 blocks of equal size <= 16), blocks numbered by
// their least site; for each candidate partner b of site 0 the finest partition invariant under the
// elements that joins 0 and b (union-find closed under the permutations).
std::vector<int> find_blocks(const std::vector<std::vector<int>>& sp, int N, int& L) {
    L = 0;
    std::vector<int> best;
    std::vector<int> parent(N);
    auto root = [&parent](int x) {
        while (parent[x] != x) {
            parent[x] = parent[parent[x]];
            x = parent[x];
        }
        return x;
    };
    for (int b = 1; b < N; ++b) {
        std::iota(parent.begin(), parent.end(), 0);
        parent[b] = 0;
        std::vector<std::pair<int, int>> todo{{0, b}};
        while (!todo.empty()) {
            const auto [x, y] = todo.back();
            todo.pop_back();
            for (const auto& p : sp) {
                const int px = p[x], py = p[y];
                const int u = root(px), v = root(py);
                if (u == v) continue;
                parent[std::max(u, v)] = std::min(u, v);
                todo.emplace_back(px, py);
            }
        }
        std::vector<int> size(N, 0);
        for (int s = 0; s < N; ++s) ++size[root(s)];
        const int l = size[0];
        if (l > 16 || l <= L || N % l != 0 || N / l < 2 || N / l > 8) continue;
        bool equal = true;
        for (int s = 0; s < N && equal; ++s) equal = size[root(s)] == l;
        if (!equal) continue;
        std::vector<int> root_block(N, -1), block(N);
        int nb = 0;
        for (int s = 0; s < N; ++s) {
            int& rb = root_block[root(s)];
            if (rb < 0) rb = nb++;
            block[s] = rb;
        }
        L = l;
        best = std::move(block);
    }
    return best;
}

}  // namespace

SublatticeTables build_sublattice(const std::vector<int>& perms, const std::vector<std::uint64_t>& flips, int G,
                                  int N, int mode, bool relaxed) {
    SublatticeTables c;
    if (G <= 1 || N < 4 || N > 64 || G > 65535 || mode == 0) return c;
    std::vector<std::vector<int>> sp;
    for (int g = 0; g < G; ++g) {
        std::vector<int> p(perms.begin() + static_cast<std::ptrdiff_t>(g) * N,
                           perms.begin() + static_cast<std::ptrdiff_t>(g + 1) * N);
        if (std::find(sp.begin(), sp.end(), p) == sp.end()) sp.push_back(std::move(p));
    }
    if (mode < 0) {
        const std::size_t min_perms = relaxed ? 16 : 64;
        if (N < 24 || sp.size() < min_perms) return c;
    }
    int L = 0;
    const std::vector<int> block = find_blocks(sp, N, L);
    if (block.empty()) return c;
    c.n = N;
    c.L = L;
    c.m = N / L;
    c.bpw = (N + 7) / 8;
    c.key_bit.assign(N, 0);
    std::vector<std::vector<int>> sites(c.m);
    for (int s = 0; s < N; ++s) sites[block[s]].push_back(s);
    for (int j = 0; j < c.m; ++j)
        for (int t = 0; t < L; ++t) c.key_bit[sites[j][t]] = N - (j + 1) * L + t;
    c.to_key.assign(static_cast<std::size_t>(c.bpw) * 256, 0ULL);
    for (int b = 0; b < c.bpw; ++b)
        for (int v = 0; v < 256; ++v) {
            std::uint64_t o = 0;
            for (int t = 0; t < 8 && 8 * b + t < N; ++t)
                if ((v >> t) & 1) o |= 1ULL << c.key_bit[8 * b + t];
            c.to_key[static_cast<std::size_t>(b) * 256 + v] = o;
        }
    // Element g fills the leading block's t-th key bit (block 0's t-th site i) from source site
    // perms[g][i]; those sources form one block src(g), so the leading block of g's image is an OR
    // of per-bit masks e_g over the pattern of block src(g), XOR g's flip bits there.
    const std::size_t P = std::size_t{1} << L;
    const std::uint32_t lmask = (L >= 32) ? 0xFFFFFFFFu : ((1u << L) - 1u);
    std::vector<int> pos(N, -1);
    for (int j = 0; j < c.m; ++j)
        for (int t = 0; t < L; ++t) pos[sites[j][t]] = t;
    std::vector<int> src(G);
    std::vector<std::vector<std::uint32_t>> e(G, std::vector<std::uint32_t>(L, 0u));
    std::vector<std::uint32_t> fl(G, 0u);
    for (int g = 0; g < G; ++g) {
        const int* p = perms.data() + static_cast<std::size_t>(g) * N;
        src[g] = block[p[sites[0][0]]];
        for (int t = 0; t < L; ++t) {
            const int i = sites[0][t];
            e[g][pos[p[i]]] = 1u << t;
            if ((flips[g] >> i) & 1ULL) fl[g] |= 1u << t;
        }
    }
    const std::size_t M = static_cast<std::size_t>(c.m) * P;
    c.lead.assign(M, 0xFFFFu);
    std::vector<std::uint32_t> img(P), count(M, 0);
    auto images = [&](int g) {
        img[0] = 0;
        for (std::size_t x = 1; x < P; ++x) img[x] = img[x & (x - 1)] | e[g][__builtin_ctzll(x)];
        for (std::size_t x = 0; x < P; ++x) img[x] = (img[x] ^ fl[g]) & lmask;
    };
    for (int g = 0; g < G; ++g) {
        images(g);
        std::uint16_t* row = c.lead.data() + static_cast<std::size_t>(src[g]) * P;
        for (std::size_t x = 0; x < P; ++x)
            if (img[x] < row[x]) row[x] = static_cast<std::uint16_t>(img[x]);
    }
    for (int g = 0; g < G; ++g) {
        images(g);
        const std::size_t base = static_cast<std::size_t>(src[g]) * P;
        for (std::size_t x = 0; x < P; ++x) count[base + x] += c.lead[base + x] == img[x] ? 1u : 0u;
    }
    c.cand_off.assign(M + 1, 0);
    for (std::size_t i = 0; i < M; ++i) c.cand_off[i + 1] = c.cand_off[i] + count[i];
    c.cand.assign(c.cand_off[M], 0);
    std::vector<std::uint32_t> fill(c.cand_off.begin(), c.cand_off.end() - 1);
    for (int g = 0; g < G; ++g) {
        images(g);
        const std::size_t base = static_cast<std::size_t>(src[g]) * P;
        for (std::size_t x = 0; x < P; ++x)
            if (c.lead[base + x] == img[x]) c.cand[fill[base + x]++] = static_cast<std::uint16_t>(g);
    }
    return c;
}

}  // namespace qsa
