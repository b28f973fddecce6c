// Spin-1/2 operator algebra and program compilation. The arithmetic follows the reference's
// src/ops/algebra.cpp and src/ops/program.cpp step for step (the same canonical expansion, the same
// coefficient products and the same key order), so a program built here holds the reference's
// coefficients bit for bit; the data structures are this project's own.
#include "algebra.h"

#include <map>

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace qsa {

namespace {
inline int popc(std::uint64_t x) { return __builtin_popcountll(x); }
inline std::uint64_t lowest(std::uint64_t x) { return x & (~x + 1ULL); }
inline std::uint64_t all_bits(int n) { return n >= 64 ? ~0ULL : ((1ULL << n) - 1ULL); }
inline std::uint64_t permute_mask(std::uint64_t m, const int* perm, int n) {
    std::uint64_t r = 0;
    for (int i = 0; i < n; ++i) r |= ((m >> perm[i]) & 1ULL) << i;
    return r;
}
}  // namespace

SpinOperator::SpinOperator(int n_sites) : n_(n_sites) {
    if (n_sites < 1 || n_sites > 64) throw std::invalid_argument("SpinOperator: n_sites must be in [1, 64]");
}

void SpinOperator::accumulate(const Key& k, cplx c) {
    if (c == cplx(0.0, 0.0)) return;
    auto it = terms_.try_emplace(k, cplx(0.0, 0.0)).first;
    it->second += c;
    if (it->second == cplx(0.0, 0.0)) terms_.erase(it);
}

// Canonical expansion (reference src/ops/algebra.cpp:34-65): a Z on a conditioned site is a known
// sign; a flip without condition splits into S+ + S-; a condition without flip is (1 +- Z)/2.
void SpinOperator::add_general(std::uint64_t C, std::uint64_t V, std::uint64_t F, std::uint64_t S, cplx c) {
    if (((C | F | S) & ~all_bits(n_)) != 0) throw std::invalid_argument("SpinOperator: mask beyond n_sites");
    struct Item {
        std::uint64_t C, V, F, S;
        cplx c;
    };
    std::vector<Item> stack{{C, V & C, F, S, c}};
    while (!stack.empty()) {
        Item it = stack.back();
        stack.pop_back();
        it.V &= it.C;
        if (const std::uint64_t known = it.S & it.C) {
            if (popc(known & it.V) & 1) it.c = -it.c;
            it.S &= ~it.C;
        }
        if (it.c == cplx(0.0, 0.0)) continue;
        if (const std::uint64_t free_flip = it.F & ~it.C) {
            const std::uint64_t b = lowest(free_flip);
            stack.push_back({it.C | b, it.V, it.F, it.S, it.c});
            stack.push_back({it.C | b, it.V | b, it.F, it.S, it.c});
            continue;
        }
        if (const std::uint64_t proj = it.C & ~it.F) {
            const std::uint64_t b = lowest(proj);
            const double s = (it.V & b) ? -1.0 : 1.0;
            stack.push_back({it.C & ~b, it.V & ~b, it.F, it.S, 0.5 * it.c});
            stack.push_back({it.C & ~b, it.V & ~b, it.F, it.S | b, 0.5 * s * it.c});
            continue;
        }
        accumulate(Key{it.F, it.V, it.S}, it.c);
    }
}

SpinOperator SpinOperator::product(int n_sites, const std::string& ops, const std::vector<int>& sites, cplx coeff) {
    if (ops.size() != sites.size()) throw std::invalid_argument("SpinOperator::product: ops and sites differ in length");
    SpinOperator r(n_sites);
    r.accumulate(Key{0, 0, 0}, coeff);
    for (std::size_t k = 0; k < ops.size(); ++k) {
        const int i = sites[k];
        if (i < 0 || i >= n_sites) throw std::invalid_argument("SpinOperator::product: site out of range");
        const std::uint64_t b = 1ULL << i;
        SpinOperator f(n_sites);
        // A set bit is spin up; Z = (-1)^bit, so S^z = -Z/2.
        switch (ops[k]) {
        case '+': f.add_general(b, 0, b, 0, 1.0); break;
        case '-': f.add_general(b, b, b, 0, 1.0); break;
        case 'z': f.add_general(0, 0, 0, b, -0.5); break;
        case 'x':
            f.add_general(b, 0, b, 0, 0.5);
            f.add_general(b, b, b, 0, 0.5);
            break;
        case 'y':
            f.add_general(b, 0, b, 0, cplx(0.0, -0.5));
            f.add_general(b, b, b, 0, cplx(0.0, 0.5));
            break;
        case 'u': f.add_general(b, b, 0, 0, 1.0); break;
        case 'd': f.add_general(b, 0, 0, 0, 1.0); break;
        case 'I': f.accumulate(Key{0, 0, 0}, 1.0); break;
        default: throw std::invalid_argument(std::string("SpinOperator::product: unknown op '") + ops[k] + "'");
        }
        r = r * f;
    }
    return r;
}

SpinOperator& SpinOperator::add(const SpinOperator& o, cplx scale) {
    if (o.n_ != n_) throw std::invalid_argument("SpinOperator::add: n_sites differ");
    if (&o == this) {
        SpinOperator r(n_);
        for (const auto& [k, c] : terms_) r.accumulate(k, (cplx(1.0, 0.0) + scale) * c);
        *this = r;
        return *this;
    }
    for (const auto& [k, c] : o.terms_) accumulate(k, scale * c);
    return *this;
}

// (src/ops/algebra.cpp:131-150): A after B; A's condition seen on the input is V_A ^ (F_B & F_A).
SpinOperator SpinOperator::operator*(const SpinOperator& o) const {
    if (o.n_ != n_) throw std::invalid_argument("SpinOperator::operator*: n_sites differ");
    SpinOperator r(n_);
    for (const auto& [ka, ca] : terms_) {
        const auto [FA, VA, SA] = ka;
        for (const auto& [kb, cb] : o.terms_) {
            const auto [FB, VB, SB] = kb;
            const std::uint64_t VA_in = VA ^ (FB & FA);
            const std::uint64_t overlap = FA & FB;
            if (((VB ^ VA_in) & overlap) != 0) continue;
            cplx c = ca * cb;
            if (popc(FB & SA) & 1) c = -c;
            r.add_general(FA | FB, VB | VA_in, FA ^ FB, SA ^ SB, c);
        }
    }
    return r;
}

SpinOperator SpinOperator::dagger() const {
    SpinOperator r(n_);
    for (const auto& [k, c] : terms_) {
        const auto [F, V, S] = k;
        r.accumulate(Key{F, V ^ F, S}, std::conj(c));
    }
    return r;
}

SpinOperator SpinOperator::image(const int* perm, std::uint64_t m) const {
    SpinOperator r(n_);
    for (const auto& [k, c] : terms_) {
        const auto [F, V, S] = k;
        const std::uint64_t Fp = permute_mask(F, perm, n_);
        const std::uint64_t Vp = permute_mask(V, perm, n_) ^ (m & Fp);
        const std::uint64_t Sp = permute_mask(S, perm, n_);
        r.accumulate(Key{Fp, Vp, Sp}, (popc(m & Sp) & 1) ? -c : c);
    }
    return r;
}

std::vector<Term> SpinOperator::terms(double drop) const {
    std::vector<Term> out;
    out.reserve(terms_.size());
    for (const auto& [k, c] : terms_) {
        if (!(std::abs(c) > drop)) continue;
        const auto [F, V, S] = k;
        out.push_back({F, V, S, c});
    }
    return out;
}

double SpinOperator::max_abs() const {
    double m = 0.0;
    for (const auto& [k, c] : terms_) m = std::max(m, std::abs(c));
    return m;
}

double SpinOperator::l1_norm() const {
    double s = 0.0;
    for (const auto& [k, c] : terms_) s += std::abs(c);
    return s;
}

bool SpinOperator::is_real(double rtol) const {
    const double scale = max_abs();
    for (const auto& [k, c] : terms_)
        if (std::abs(c.imag()) > rtol * scale) return false;
    return true;
}

int SpinOperator::delta_set_bits() const {
    bool first = true;
    int delta = 0;
    for (const auto& [k, c] : terms_) {
        const auto [F, V, S] = k;
        const int d = popc(F & ~V) - popc(F & V);
        if (first) {
            delta = d;
            first = false;
        } else if (d != delta) {
            throw std::invalid_argument("SpinOperator::delta_set_bits: terms change S^z by different amounts");
        }
    }
    return delta;
}

namespace {
struct Packed {
    std::uint64_t sign;
    cplx c;
};
using Tree = std::map<std::uint64_t, std::map<std::uint64_t, std::vector<Packed>>>;

Program pack(const Tree& tree) {
    Program P;
    P.group_vbegin.push_back(0);
    P.vsub_tbegin.push_back(0);
    for (const auto& [F, vmap] : tree) {
        P.group_flip.push_back(F);
        int setbits = popc(vmap.begin()->first);
        for (const auto& [V, ts] : vmap)
            if (popc(V) != setbits) setbits = -1;
        P.group_setbits.push_back(setbits);
        for (const auto& [V, ts] : vmap) {
            P.vsub_val.push_back(V);
            for (const auto& t : ts) {
                P.term_sign.push_back(t.sign);
                P.term_coeff.push_back(t.c);
            }
            P.vsub_tbegin.push_back(static_cast<std::uint32_t>(P.term_sign.size()));
        }
        P.group_vbegin.push_back(static_cast<std::uint32_t>(P.vsub_val.size()));
    }
    return P;
}
}  // namespace

Program compile_operator(const SpinOperator& O) {
    Tree tree;
    for (const auto& t : O.terms()) tree[t.flip][t.value].push_back({t.sign, t.coeff});
    return pack(tree);
}

Program compile_projected(const SpinOperator& O, int G, const int* perms, const std::uint64_t* flips,
                          const cplx* chi_src, const cplx* chi_tgt, int src_nup, int tgt_nup) {
    const int n = O.n_sites();
    if ((src_nup < 0) != (tgt_nup < 0))
        throw std::invalid_argument("compile_projected: a fixed-S^z sector paired with a full sector");
    const bool fixed_sz = src_nup >= 0;
    const int delta_req = fixed_sz ? tgt_nup - src_nup : 0;
    SpinOperator Ol(n);
    for (int g = 0; g < G; ++g) {
        const cplx wg = std::conj(chi_tgt[g] * std::conj(chi_src[g])) / static_cast<double>(G);
        if (std::abs(wg) < 1e-15) continue;
        Ol.add(O.image(perms + static_cast<std::size_t>(g) * n, flips ? flips[g] : 0ULL), wg);
    }
    const double scale = O.max_abs();
    Tree tree;
    for (const auto& t : Ol.terms(1e-14 * scale)) {
        if (fixed_sz) {
            const int d = popc(t.flip & ~t.value) - popc(t.flip & t.value);
            if (d != delta_req) continue;
        }
        tree[t.flip][t.value].push_back({t.sign, t.coeff});
    }
    if (tree.empty()) return Program{{}, {}, {0}, {}, {0}, {}, {}};
    return pack(tree);
}

MultiProgram merge_programs(const std::vector<const Program*>& programs) {
    struct Term {
        std::uint32_t obs;
        std::uint64_t sign;
        cplx coeff;
    };
    std::map<std::uint64_t, std::map<std::uint64_t, std::vector<Term>>> groups;
    for (std::size_t o = 0; o < programs.size(); ++o) {
        const Program& P = *programs[o];
        for (std::size_t g = 0; g < P.n_groups(); ++g)
            for (std::uint32_t v = P.group_vbegin[g]; v < P.group_vbegin[g + 1]; ++v)
                for (std::uint32_t k = P.vsub_tbegin[v]; k < P.vsub_tbegin[v + 1]; ++k)
                    groups[P.group_flip[g]][P.vsub_val[v]].push_back(
                        Term{static_cast<std::uint32_t>(o), P.term_sign[k], P.term_coeff[k]});
    }
    MultiProgram M;
    M.n_obs = static_cast<int>(programs.size());
    Program& P = M.prog;
    for (const auto& [flip, vals] : groups) {
        P.group_flip.push_back(flip);
        P.group_vbegin.push_back(static_cast<std::uint32_t>(P.vsub_val.size()));
        int setbits = -2;
        for (const auto& [val, terms] : vals) {
            const int pc = __builtin_popcountll(val);
            setbits = setbits == -2 ? pc : (setbits == pc ? setbits : -1);
            P.vsub_val.push_back(val);
            P.vsub_tbegin.push_back(static_cast<std::uint32_t>(P.term_sign.size()));
            for (const Term& t : terms) {
                P.term_sign.push_back(t.sign);
                P.term_coeff.push_back(t.coeff);
                M.term_obs.push_back(t.obs);
            }
        }
        P.group_setbits.push_back(setbits < 0 ? -1 : setbits);
    }
    P.group_vbegin.push_back(static_cast<std::uint32_t>(P.vsub_val.size()));
    P.vsub_tbegin.push_back(static_cast<std::uint32_t>(P.term_sign.size()));
    return M;
}

}  // namespace qsa
