// Operator algebra of spin-1/2 terms and the flip-grouped program the sector walk reads.
//
// A term acts on a basis state s (set bit = spin up) as
//     T|s> = c (-1)^{popcount(s & S)} |s ^ F>   if (s & F) == V, else 0,
// the canonical form of include/ed/ops/term.h:42-65 (condition only on flipped bits, Z factors on
// the others). Terms are keyed (F, V, S) and kept in ascending key order, so the expansion of an
// operator is unique and the program built from it lists terms exactly as the reference's
// compile_operator (src/ops/program.cpp:34-56, :132-141) does: groups by ascending flip mask,
// value subgroups ascending, terms of a subgroup in ascending sign mask.
#pragma once

#include <complex>
#include <cstdint>
#include <map>
#include <string>
#include <tuple>
#include <vector>

namespace qsa {

using cplx = std::complex<double>;

struct Term {
    std::uint64_t flip = 0, value = 0, sign = 0;
    cplx coeff{0.0, 0.0};
};

class SpinOperator {
public:
    explicit SpinOperator(int n_sites);

    /// coeff * ops[0](sites[0]) * ops[1](sites[1]) * ..., the last factor acting first;
    /// ops from "+-zxyudI" (S+, S-, S^z, S^x, S^y, |up><up|, |dn><dn|, identity).
    /// Follows MaskedOperator::product (src/ops/algebra.cpp:67-100).
    static SpinOperator product(int n_sites, const std::string& ops, const std::vector<int>& sites, cplx coeff);

    int n_sites() const { return n_; }
    std::size_t size() const { return terms_.size(); }

    SpinOperator& add(const SpinOperator& o, cplx scale = cplx(1.0, 0.0));
    SpinOperator operator*(const SpinOperator& o) const;   // this after o
    SpinOperator dagger() const;
    /// U_g O U_g^dagger for the element (perm in the "bit i of g s = bit perm[i] of s" convention,
    /// then XOR flip) -- src/ops/algebra.cpp:178-188.
    SpinOperator image(const int* perm, std::uint64_t flip) const;
    /// Add one term in any (cond, val, flip, sign) form, expanded to canonical terms
    /// (src/ops/algebra.cpp:34-65).
    void add_general(std::uint64_t cond, std::uint64_t val, std::uint64_t flip, std::uint64_t sign, cplx c);

    std::vector<Term> terms(double drop = 0.0) const;
    double max_abs() const;
    double l1_norm() const;   ///< s_H: the sum of |c| (an upper bound of the operator norm)
    bool is_real(double rtol = 1e-12) const;
    int delta_set_bits() const;   ///< uniform change of the number of set bits; throws if mixed

private:
    using Key = std::tuple<std::uint64_t, std::uint64_t, std::uint64_t>;
    int n_;
    std::map<Key, cplx> terms_;
    void accumulate(const Key& k, cplx c);
};

/// The flat program of one operator (the layout of include/ed/ops/program.h:51-94).
struct Program {
    std::vector<std::uint64_t> group_flip;
    std::vector<std::int32_t> group_setbits;   ///< common popcount of the group's values, or -1
    std::vector<std::uint32_t> group_vbegin;   ///< n_groups + 1
    std::vector<std::uint64_t> vsub_val;
    std::vector<std::uint32_t> vsub_tbegin;    ///< n_vsub + 1
    std::vector<std::uint64_t> term_sign;
    std::vector<cplx> term_coeff;
    std::size_t n_groups() const { return group_flip.size(); }
    std::size_t n_terms() const { return term_sign.size(); }
};

/// The program of O's canonical terms (src/ops/program.cpp:132-141). A row walk reads the
/// program of O^dagger and conjugates (include/ed/ops/operator.h:132-138).
Program compile_operator(const SpinOperator& O);

/// The lambda-projected program of `ops` for <bra|O|ket>, ket in the sector with characters
/// chi_src and n_up src_nup, bra in chi_tgt / tgt_nup, both of one group (perms row-major |G| x N,
/// flips): O_lambda = (1/|G|) sum_g conj(lambda(g)) U_g O U_g^dag, lambda = chi_tgt conj(chi_src);
/// terms below 1e-14 of the unprojected max dropped; between fixed-Sz sectors only terms with the
/// right change of set bits kept (src/ops/program.cpp:61-130).
Program compile_projected(const SpinOperator& O, int G, const int* perms, const std::uint64_t* flips,
                          const cplx* chi_src, const cplx* chi_tgt, int src_nup, int tgt_nup);

/// Several programs (each of one observable, compiled for one pair of sectors) as one: groups by
/// flip mask ascending, value subgroups ascending, and in a subgroup every observable's terms in
/// observable order, each tagged with its observable (the multi-observable MaskedProgram of the
/// reference's compile_program, include/ed/ops/program.h:60-120).
struct MultiProgram {
    Program prog;
    std::vector<std::uint32_t> term_obs;
    int n_obs = 0;
};
MultiProgram merge_programs(const std::vector<const Program*>& programs);

}  // namespace qsa
