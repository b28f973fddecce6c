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

std::vector<std::uint64_t> GroupSpec::nibble_tables() const {
    const int N = n_sites, Q = nq();
    const bool keyed = slc.engaged();
    auto keyed_bit = [&](int site) { return keyed ? slc.key_bit[site] : site; };
    std::vector<std::uint64_t> t(static_cast<std::size_t>(G) * Q * 16 + G + (keyed ? 2 * Q * 16 : 0), 0ULL);
    for (int g = 0; g < G; ++g) {
        const int* p = perms.data() + static_cast<std::size_t>(g) * N;
        int inv[64];
        for (int i = 0; i < N; ++i) inv[p[i]] = i;   // source bit j lands on output bit inv[j]
        for (int q = 0; q < Q; ++q)
            for (int v = 0; v < 16; ++v) {
                std::uint64_t o = 0;
                for (int u = 0; u < 4 && 4 * q + u < N; ++u)
                    if ((v >> u) & 1) o |= 1ULL << keyed_bit(inv[4 * q + u]);
                t[(static_cast<std::size_t>(g) * Q + q) * 16 + v] = o;
            }
        std::uint64_t f = 0;
        for (int i = 0; i < N; ++i)
            if ((flips[g] >> i) & 1ULL) f |= 1ULL << keyed_bit(i);
        t[static_cast<std::size_t>(G) * Q * 16 + g] = f;
    }
    if (keyed) {
        std::uint64_t* key = t.data() + static_cast<std::size_t>(G) * Q * 16 + G;
        std::uint64_t* unkey = key + Q * 16;
        int site_of[64];
        for (int i = 0; i < N; ++i) site_of[slc.key_bit[i]] = i;
        for (int q = 0; q < Q; ++q)
            for (int v = 0; v < 16; ++v) {
                std::uint64_t k = 0, s = 0;
                for (int u = 0; u < 4 && 4 * q + u < N; ++u)
                    if ((v >> u) & 1) {
                        k |= 1ULL << slc.key_bit[4 * q + u];
                        s |= 1ULL << site_of[4 * q + u];
                    }
                key[q * 16 + v] = k;
                unkey[q * 16 + v] = s;
            }
    }
    return t;
}

bool GroupSpec::real_characters() const {
    for (const auto& c : chars)
        if (c.imag() != 0.0) return false;
    return true;
}

namespace {

/// Disjoint sets over the sites; every set is named by its least site.
class SiteSets {
public:
    explicit SiteSets(int n) : up_(static_cast<std::size_t>(n)) { std::iota(up_.begin(), up_.end(), 0); }
    int name(int x) {
        int r = x;
        while (up_[r] != r) r = up_[r];
        while (up_[x] != r) {   // point the walked chain straight at its name
            const int next = up_[x];
            up_[x] = r;
            x = next;
        }
        return r;
    }
    void join(int a, int b) {
        a = name(a);
        b = name(b);
        if (a != b) up_[std::max(a, b)] = std::min(a, b);
    }

private:
    std::vector<int> up_;
};

/// The least block system in which sites 0 and b share a block: the connected components of the
/// orbital graph of the pair {0, b}, whose edges are its images {p[0], p[b]} under every element.
/// Because `elems` is a whole group (closed under products and inverses) this edge set is mapped
/// onto itself by every element, so its components form an invariant partition, and every invariant
/// partition joining 0 and b contains all the edges. Returns each site's block named by its least
/// site.
std::vector<int> orbital_components(const std::vector<const int*>& elems, int N, int b) {
    SiteSets sets(N);
    for (const int* p : elems) sets.join(p[0], p[b]);
    std::vector<int> name(static_cast<std::size_t>(N));
    for (int s = 0; s < N; ++s) name[s] = sets.name(s);
    return name;
}

/// The block system with the largest blocks of the admissible shapes (blocks of equal size L <= 16,
/// 2..8 of them), the first partner b of site 0 reaching that size; blocks numbered in the order of
/// their least sites. Empty when none is admissible.
std::vector<int> widest_block_system(const std::vector<const int*>& elems, int N, int& L) {
    L = 0;
    std::vector<int> chosen;
    for (int b = 1; b < N; ++b) {
        const std::vector<int> name = orbital_components(elems, N, b);
        std::vector<int> size(static_cast<std::size_t>(N), 0);
        for (int s = 0; s < N; ++s) ++size[name[s]];
        const int l = size[0];
        if (l <= L || l > 16 || N % l != 0 || N / l < 2 || N / l > 8) continue;
        if (std::any_of(name.begin(), name.end(), [&](int r) { return size[r] != l; })) continue;
        // block numbers: a block's least site is its name, and names appear first at themselves
        std::vector<int> number(static_cast<std::size_t>(N), -1), block(static_cast<std::size_t>(N));
        int next = 0;
        for (int s = 0; s < N; ++s) {
            if (name[s] == s) number[s] = next++;
            block[s] = number[name[s]];
        }
        L = l;
        chosen = std::move(block);
    }
    return chosen;
}

}  // namespace

SublatticeTables build_sublattice(const std::vector<int>& perms, const std::vector<std::uint64_t>& flips, int G,
                                  int N, int mode, bool relaxed) {
    SublatticeTables c;
    if (G <= 1 || N < 4 || N > 64 || G > 65535 || mode == 0) return c;
    // the distinct site permutations (the flip half of a group repeats them)
    std::vector<const int*> elems;
    for (int g = 0; g < G; ++g) {
        const int* p = perms.data() + static_cast<std::size_t>(g) * N;
        const bool seen = std::any_of(elems.begin(), elems.end(), [&](const int* q) { return std::equal(p, p + N, q); });
        if (!seen) elems.push_back(p);
    }
    if (mode < 0) {
        const std::size_t min_perms = relaxed ? 16 : 64;
        if (N < 24 || elems.size() < min_perms) return c;
    }
    int L = 0;
    const std::vector<int> block = widest_block_system(elems, N, L);
    if (block.empty()) return c;
    c.n = N;
    c.L = L;
    c.m = N / L;
    c.bpw = (N + 7) / 8;

    // block j holds the sites `members[j]` ascending; its t-th site sits at key bit N - (j+1) L + t
    std::vector<std::vector<int>> members(static_cast<std::size_t>(c.m));
    std::vector<int> slot(static_cast<std::size_t>(N));
    for (int s = 0; s < N; ++s) {
        slot[s] = static_cast<int>(members[block[s]].size());
        members[block[s]].push_back(s);
    }
    c.key_bit.resize(static_cast<std::size_t>(N));
    for (int s = 0; s < N; ++s) c.key_bit[s] = N - (block[s] + 1) * L + slot[s];
    c.to_key.assign(static_cast<std::size_t>(c.bpw) * 256, 0ULL);
    for (int byte = 0; byte < c.bpw; ++byte)
        for (int v = 1; v < 256; ++v) {
            const int low = __builtin_ctz(static_cast<unsigned>(v));
            const std::uint64_t rest = c.to_key[static_cast<std::size_t>(byte) * 256 + (v & (v - 1))];
            const int site = 8 * byte + low;
            c.to_key[static_cast<std::size_t>(byte) * 256 + v] = site < N ? rest | (1ULL << c.key_bit[site]) : rest;
        }

    // The leading block of an image g s (key bits [N - L, N)) is read off one block of s: its t-th
    // bit is bit p[members[0][t]] of s, XOR g's flip there, and those sites all lie in one block,
    // feed(g). Per element, the image of a block pattern x is two byte lookups and the flip.
    const int P = 1 << L;
    struct Lead {
        int feed = 0;
        std::uint16_t flip = 0;
        std::uint16_t byte_img[2][256];
    };
    std::vector<Lead> lead_of(static_cast<std::size_t>(G));
    for (int g = 0; g < G; ++g) {
        const int* p = perms.data() + static_cast<std::size_t>(g) * N;
        Lead& e = lead_of[g];
        e.feed = block[p[members[0][0]]];
        int to_bit[16];   // slot u of block feed(g) -> bit t of the leading pattern
        for (int t = 0; t < L; ++t) {
            const int site = p[members[0][t]];
            if (block[site] != e.feed) throw std::logic_error("sublattice: an element splits a block");
            to_bit[slot[site]] = t;
            if ((flips[g] >> members[0][t]) & 1ULL) e.flip = static_cast<std::uint16_t>(e.flip | (1u << t));
        }
        for (int half = 0; half < 2; ++half)
            for (int v = 0; v < 256; ++v) {
                std::uint32_t o = 0;
                for (int u = 0; u < 8; ++u)
                    if ((v >> u) & 1 && 8 * half + u < L) o |= 1u << to_bit[8 * half + u];
                e.byte_img[half][v] = static_cast<std::uint16_t>(o);
            }
    }
    // the elements fed by each block, ascending
    std::vector<std::vector<int>> fed_by(static_cast<std::size_t>(c.m));
    for (int g = 0; g < G; ++g) fed_by[lead_of[g].feed].push_back(g);

    // Entry (j, x): the least leading pattern any element fed by block j makes of pattern x, and
    // the elements that make it, ascending. Entries are filled in order, so the candidate lists
    // are appended one after another.
    const std::size_t M = static_cast<std::size_t>(c.m) * static_cast<std::size_t>(P);
    c.lead.assign(M, 0xFFFFu);
    c.cand_off.assign(M + 1, 0);
    c.cand.clear();
    std::vector<std::uint16_t> img;
    for (int j = 0; j < c.m; ++j) {
        const std::vector<int>& E = fed_by[j];
        img.resize(E.size());
        for (int x = 0; x < P; ++x) {
            const std::size_t entry = static_cast<std::size_t>(j) * P + static_cast<std::size_t>(x);
            std::uint16_t least = 0xFFFFu;
            for (std::size_t i = 0; i < E.size(); ++i) {
                const Lead& e = lead_of[E[i]];
                img[i] = static_cast<std::uint16_t>((e.byte_img[0][x & 255] | e.byte_img[1][x >> 8]) ^ e.flip);
                least = std::min(least, img[i]);
            }
            c.lead[entry] = least;
            for (std::size_t i = 0; i < E.size(); ++i)
                if (img[i] == least) c.cand.push_back(static_cast<std::uint16_t>(E[i]));
            c.cand_off[entry + 1] = static_cast<std::uint32_t>(c.cand.size());
        }
    }
    return c;
}

}  // namespace qsa
