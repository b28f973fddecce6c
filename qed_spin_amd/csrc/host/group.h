// The symmetry group of a sector in the form the device kernels read: byte lookup tables for the
// element action and, when the group permutes a block system, the sublattice key-order tables that
// decide which orbit member is the representative.
#pragma once

#include <complex>
#include <cstdint>
#include <memory>
#include <vector>

namespace qsa {

using cplx = std::complex<double>;

/// The key order of the sublattice representative search (the reference's
/// include/ed/basis/sublattice_code.h:64-400): sites in blocks of L, block j in key bits
/// [N - (j+1) L, N - j L); the representative is the orbit member least in this order.
struct SublatticeTables {
    int L = 0, m = 0, bpw = 0, n = 0;
    std::vector<int> key_bit;              ///< per site: its key position
    std::vector<std::uint64_t> to_key;     ///< bpw x 256: a state's bits moved to key positions
    std::vector<std::uint16_t> lead;       ///< m x 2^L: least leading block reachable from block j's pattern
    std::vector<std::uint32_t> cand_off;   ///< m x 2^L + 1
    std::vector<std::uint16_t> cand;       ///< elements reaching the least leading block, ascending
    bool engaged() const { return L > 0; }
};

/// Build the key-order tables, or return an empty (not engaged) object. mode: 0 never, 1 whenever a
/// block system of 2..8 blocks of at most 16 sites exists, -1 the size rule (N >= 24 and at least
/// 16 distinct permutations when `relaxed`, 64 otherwise) -- the reference's
/// SublatticeCode::of / build_ (sublattice_code.h:187-230, 286-400).
SublatticeTables build_sublattice(const std::vector<int>& perms, const std::vector<std::uint64_t>& flips, int G,
                                  int N, int mode, bool relaxed);

/// One symmetry sector's group data: |G| elements (site permutation, then XOR flip), its
/// characters, and the fixed magnetisation (n_up < 0: the full 2^N space).
struct GroupSpec {
    int n_sites = 0;
    int n_up = -1;
    int G = 0;
    std::vector<int> perms;               ///< G x N row-major: bit i of g s = bit perms[g N + i] of s
    std::vector<std::uint64_t> flips;     ///< G
    std::vector<cplx> chars;              ///< G
    SublatticeTables slc;

    int bpw() const { return (n_sites + 7) / 8; }
    int nq() const { return (n_sites + 3) / 4; }
    /// The action by nibbles, one array: the images of nibble q = v under every element's
    /// permutation ([(g nq + q) 16 + v]), then the G flips, then (sublattice order engaged) the
    /// nibbles' key positions and their inverse (nq x 16 each); images and flips in key positions
    /// when the sublattice order is engaged. For kernels that hold the action in LDS.
    std::vector<std::uint64_t> nibble_tables() const;
    /// Byte lookup tables of the action: lut[(g bpw + b) 256 + v] is the image of byte b = v
    /// (include/ed/basis/compiled_group.h:105-162 layout).
    std::vector<std::uint64_t> byte_lut() const;
    bool real_characters() const;
};

}  // namespace qsa
