"""Level vectors of the eigs verb: their expansion out of the symmetry sector, observables in them,
and the .npz file of a result -- the reference's EigResult.vectors / expect / correlations / save
and qed.load_eigs (python/qed/_verbs/eigs.py:52-245), with the array layout of its
eigs_to_arrays / eigs_from_arrays (python/qed/_bindings/sectors.cpp:56-330).

A level vector is kept in the sector basis it was solved in (one LevelBasis per block: the group,
n_up, representatives and 1/norms). The expansion into the Sz sector or the full 2^N space runs on
the device (qedh_sector_expand: one thread per representative writes its orbit). Expectation values
apply each observable's projected program between two handles of the level's sector (gather lane);
correlations <A_a^dag B_b> expand the vector into the Sz sector once and apply every A_a and B_b
there, split by the change of magnetisation of their products (parts that change n_up differently
land in different Sz sectors and are orthogonal).

The file: format_version 2, k, energies, the level arrays (level_energy, level_mirror, level_fold,
level_multiplicity, level_vector, level_tag with 11 entries per level, the ragged momentum and
irrep labels), per basis b the arrays basis{b}_reps / _inv_norms / _characters / _perms /
_flip_masks / _shape / _sublattice, and vector_basis / vector_offset / vector_amplitudes, n_bases,
spec_* and result_scalars. Every level this walk returns is its own (block, eigenvalue) entry: no
antiunitary fold, no mirror partner, multiplicity 1, star of one block. Files of sectors of irreps of
dimension above 1 are refused. Every index and offset of a file is checked before it is used.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _lib
from .errors import InvalidRequest
from .operator import Operator
from .sector import LANE_GATHER, DeviceArray, GroupSpec, Sector, SectorMatvec, matrix_elements, projected_program
from .symmetry import trivial_group

FORMAT_VERSION = 2
TAG_WIDTH = 11     # n_up, sz_parity, k0, k_raw, flip_parity, irrep, irrep_dim, star_size, tr_folded, dim, multiplicity
_MASK64 = (1 << 64) - 1


def sublattice_fingerprint(key_bits) -> int:
    """The 64-bit FNV-1a hash (offset basis 0xcbf29ce484222325, prime 0x100000001b3) of key_bit[s] + 1
    over the sites in order: the value a saved basis carries to name the sublattice key order its
    representatives were chosen in (include/ed/basis/sublattice_code.h:326-332); 0 names the plain
    order."""
    h = 0xCBF29CE484222325
    for kb in key_bits:
        h ^= (int(kb) + 1) & _MASK64
        h = (h * 0x100000001B3) & _MASK64
    return h


@dataclass
class LevelBasis:
    """The sector a level vector lives in (the reference's RepSectorData of a BlockVector)."""

    perms: np.ndarray          # (G, N) int32
    chars: np.ndarray          # (G,) complex128
    flips: np.ndarray          # (G,) uint64
    n_up: int
    reps: np.ndarray           # (dim,) uint64, ascending
    inv_norms: np.ndarray      # (dim,) float64
    sublattice: int = 0        # fingerprint of the key order (0: plain order)
    _sector: Optional[Sector] = field(default=None, repr=False, compare=False)

    @classmethod
    def of(cls, group: GroupSpec, sec: Sector) -> "LevelBasis":
        reps, inv = sec.export()
        kb, L = sec.key_bits()
        flips = np.zeros(group.G, np.uint64) if group.flips is None else np.asarray(group.flips, np.uint64)
        b = cls(perms=np.asarray(group.perms, np.int32), chars=np.asarray(group.chars, np.complex128), flips=flips,
                n_up=int(sec.n_up), reps=reps, inv_norms=inv,
                sublattice=sublattice_fingerprint(kb) if L else 0)
        b._sector = sec
        return b

    @property
    def N(self) -> int:
        return int(self.perms.shape[1])

    @property
    def dim(self) -> int:
        return int(self.reps.size)

    def group(self) -> GroupSpec:
        flips = self.flips if np.any(self.flips) else None
        return GroupSpec(perms=self.perms, chars=self.chars, flips=flips)

    def sector(self) -> Sector:
        """The device sector of these representatives (imported once: no enumeration; each
        representative is checked to find itself through the group's lookup)."""
        if self._sector is None:
            s = Sector(self.group(), self.n_up, sublattice=1 if self.sublattice else 0, reps=self.reps,
                       inv_norms=self.inv_norms)
            if self.sublattice:
                kb, L = s.key_bits()
                if not L or sublattice_fingerprint(kb) != self.sublattice:
                    raise InvalidRequest("load_eigs: a basis was computed with a sublattice key order this "
                                         "build does not reproduce")
            self._sector = s
        return self._sector


def _sz_dim(N: int, n_up: int) -> int:
    from math import comb
    return comb(N, n_up)


def expand(basis: LevelBasis, vecs: Sequence[np.ndarray], full: bool) -> List[np.ndarray]:
    """The sector vectors as unit vectors of the Sz sector of basis.n_up (states ascending) or of
    the full 2^N space, expanded on the device."""
    if not vecs:
        return []
    N = basis.N
    if full and N > 34:
        raise InvalidRequest("vectors: the full 2^N space is limited to N <= 34 (use basis='sz')")
    if not full and basis.n_up < 0:
        raise InvalidRequest("vectors: a sector without fixed n_up expands only into the full space")
    for v in vecs:
        if np.asarray(v).size != basis.dim:
            raise InvalidRequest(f"vectors: a vector of {np.asarray(v).size} entries for a basis of {basis.dim}")
    out_dim = (1 << N) if full else _sz_dim(N, basis.n_up)
    U = DeviceArray.from_host(np.concatenate([np.asarray(v, np.complex128) for v in vecs]))
    out = DeviceArray(len(vecs) * out_dim)
    norms = np.zeros(len(vecs))
    _lib.call("qedh_sector_expand", basis.sector().h, U.ptr, len(vecs), int(bool(full)), out.ptr,
              _lib.ptr(norms, C.c_double))
    psi = out.to_host()
    return [psi[i * out_dim:(i + 1) * out_dim].copy() for i in range(len(vecs))]


def _vdot(x_ptr, y_ptr, n: int) -> complex:
    out = np.zeros(2)
    _lib.call("qedh_vdot", x_ptr, y_ptr, C.c_int64(int(n)), 1, _lib.ptr(out, C.c_double))
    return complex(out[0], out[1])


def expect_in(basis: LevelBasis, vec: np.ndarray, ops: Sequence[Operator], cache: Dict) -> np.ndarray:
    """<v|O|v> (v a unit vector of the sector) of every O in one walk of the sector (the fused
    observable sweep, qedh_matrix_elements): each O's projected program between two handles of the
    sector (an observable need not be group invariant)."""
    sec = basis.sector()
    if "twin" not in cache:
        cache["twin"] = Sector(basis.group(), basis.n_up, sublattice=1 if basis.sublattice else 0, reps=basis.reps,
                               inv_norms=basis.inv_norms)
    v = DeviceArray.from_host(np.asarray(vec, np.complex128))
    programs = []
    for O in ops:
        key = ("prog", id(O))
        if key not in cache:
            cache[key] = (O, projected_program(O, cache["twin"], sec))
        programs.append(cache[key][1])
    return matrix_elements(ops, cache["twin"], sec, [v], [v], programs=programs)[0]


def split_by_shift(op: Operator) -> Dict[int, Operator]:
    """O as a sum of parts that each change n_up by one amount d: S^x = (S^+ + S^-)/2 and
    S^y = (S^+ - S^-)/(2i) are written out, then the products grouped by (#'+' - #'-')."""
    parts: Dict[int, Operator] = {}
    for ops, sites, c in op.products:
        terms = [("", 1.0 + 0j)]
        for ch in ops:
            if ch == "x":
                choices = (("+", 0.5), ("-", 0.5))
            elif ch == "y":
                choices = (("+", -0.5j), ("-", 0.5j))
            else:
                choices = ((ch, 1.0),)
            terms = [(s + a, w * f) for s, w in terms for a, f in choices]
        for s, w in terms:
            d = s.count("+") - s.count("-")
            parts.setdefault(d, Operator(op.num_sites)).add_product(s, sites, c * w)
    return parts


def correlations_in(basis: LevelBasis, vec: np.ndarray, A: Sequence[Operator], B: Sequence[Operator],
                    cache: Dict) -> np.ndarray:
    """C[a, b] = <A_a v | B_b v> = <v| A_a^dag B_b |v> for the unit sector vector v: v expanded into
    the Sz sector (the full space for a sector without fixed n_up), each part of A_a and B_b that
    changes n_up by d applied there into the Sz sector n_up + d, and the overlaps of equal d summed."""
    N = basis.N
    fixed = basis.n_up >= 0
    psi = expand(basis, [vec], full=not fixed)[0]
    if "sz" not in cache:
        cache["sz"] = {}
    secs = cache["sz"]

    def sz_sector(n):
        if n not in secs:
            secs[n] = Sector(trivial_group(N), n)
        return secs[n]

    src = sz_sector(basis.n_up if fixed else -1)
    x = DeviceArray.from_host(psi)

    def images(ops):
        out = []
        for O in ops:
            parts = split_by_shift(O) if fixed else {0: O}
            img = {}
            for d, P in parts.items():
                n = basis.n_up + d
                if fixed and not 0 <= n <= N:
                    continue
                tgt = sz_sector(n if fixed else -1)
                if tgt.dim == 0:
                    continue
                y = DeviceArray(tgt.dim)
                SectorMatvec(P, tgt, src, lane=LANE_GATHER).apply(x, y, complex_vectors=True)
                img[d] = (y, tgt.dim)
            out.append(img)
        return out

    ia = images(A)
    ib = ia if B is A else images(B)
    Cm = np.zeros((len(A), len(B)), np.complex128)
    for a, wa in enumerate(ia):
        for b, wb in enumerate(ib):
            Cm[a, b] = sum((_vdot(wa[d][0].ptr, wb[d][0].ptr, wa[d][1]) for d in wa if d in wb), 0j)
    return Cm


def tower_members(basis: LevelBasis, vec: np.ndarray, full: bool, lowest_n_up: int, cache: Dict) -> List[np.ndarray]:
    """The members of a spin multiplet from its top member v (the sector vector at n_up = basis.n_up):
    v expanded into its Sz sector, then S^- = sum_i S^-_i applied on the device (gather lane between
    the Sz sectors n and n - 1) down to n_up = lowest_n_up, each member normalised (the reference's
    multiplet, src/engine/eigs.cpp:749-860, total S^- for an SU(2) multiplet). Returned top first, in
    the Sz sectors (states ascending) or, full, in the 2^N space."""
    N = basis.N
    psi = expand(basis, [vec], full=False)[0]
    secs = cache.setdefault("sz", {})

    def sz_sector(n):
        if n not in secs:
            secs[n] = Sector(trivial_group(N), n)
        return secs[n]

    def place(n, p):
        if not full:
            return p
        reps, _ = sz_sector(n).export()
        out = np.zeros(1 << N, np.complex128)
        out[reps.astype(np.int64)] = p
        return out

    members = [place(basis.n_up, psi)]
    lower = cache.get("S-")
    if lower is None:
        lower = Operator(N)
        for i in range(N):
            lower.add_product("-", (i,), 1.0)
        cache["S-"] = lower
    x = DeviceArray.from_host(psi)
    for n in range(basis.n_up, lowest_n_up, -1):
        tgt = sz_sector(n - 1)
        y = DeviceArray(tgt.dim)
        SectorMatvec(lower, tgt, sz_sector(n), lane=LANE_GATHER).apply(x, y, complex_vectors=True)
        p = y.to_host()
        nrm = np.linalg.norm(p)
        if not nrm > 0.0:
            raise InvalidRequest("vectors: S^- annihilated a member above the bottom of its multiplet")
        p /= nrm
        members.append(place(n - 1, p))
        x = DeviceArray.from_host(p)
    return members


# ---- the .npz file ----------------------------------------------------------------------------------

def to_arrays(result) -> Dict[str, np.ndarray]:
    """The arrays of a result (eigs_to_arrays, python/qed/_bindings/sectors.cpp:56-162)."""
    levels = result.levels
    nl = len(levels)
    d: Dict[str, np.ndarray] = {}
    d["format_version"] = np.int64(FORMAT_VERSION)
    d["k"] = np.int64(result.k)
    d["energies"] = np.asarray(result.energies, np.float64)
    d["level_energy"] = np.array([L["energy"] for L in levels], np.float64)
    d["level_mirror"] = np.full(nl, -1, np.int64)
    d["level_fold"] = np.zeros(nl, np.int64)
    d["level_multiplicity"] = np.array([L.get("multiplicity", 1) for L in levels], np.int64)
    d["level_vector"] = np.array([L.get("vector", -1) for L in levels], np.int64)
    tag = np.zeros((nl, TAG_WIDTH), np.int64)
    mchi, moff = [], [0]
    for i, L in enumerate(levels):
        m = L.get("momentum")
        par = L.get("parity")
        tag[i] = [L["n_up"], 0, 0 if m is None else m, -1 if m is None else m, 0,
                  -1 if par is None else (0 if par > 0 else 1), 1, 1, 0, L["dim"], L.get("multiplicity", 1)]
        if m is not None:
            mchi.append(np.exp(2j * np.pi * m / result.n_sites))
        moff.append(len(mchi))
    d["level_tag"] = tag.reshape(-1)
    d["level_momentum"] = np.array(mchi, np.complex128)
    d["level_momentum_offset"] = np.array(moff, np.int64)
    d["level_irrep_elems"] = np.zeros(0, np.int64)
    d["level_irrep_chars"] = np.zeros(0, np.complex128)
    d["level_irrep_offset"] = np.zeros(nl + 1, np.int64)
    bases = result._bases
    order = sorted(bases)
    bid = {b: i for i, b in enumerate(order)}
    for b in order:
        B = bases[b]
        p = f"basis{bid[b]}_"
        d[p + "reps"] = np.asarray(B.reps, np.uint64)
        d[p + "inv_norms"] = np.asarray(B.inv_norms, np.float64)
        d[p + "characters"] = np.asarray(B.chars, np.complex128)
        d[p + "perms"] = np.asarray(B.perms, np.int64).reshape(-1)
        d[p + "flip_masks"] = np.asarray(B.flips, np.uint64)
        d[p + "shape"] = np.array([B.perms.shape[0], B.N, B.n_up, 1], np.int64)
        d[p + "sublattice"] = np.array([B.sublattice], np.uint64)
    vb, voff, amps = [], [0], []
    for basis_key, v in result._vectors:
        vb.append(bid[basis_key])
        amps.append(np.asarray(v, np.complex128))
        voff.append(voff[-1] + len(v))
    d["n_bases"] = np.array([len(order)], np.int64)
    d["vector_basis"] = np.array(vb, np.int64)
    d["vector_offset"] = np.array(voff, np.int64)
    d["vector_amplitudes"] = np.concatenate(amps) if amps else np.zeros(0, np.complex128)
    d["spec_abelian"] = np.zeros(0, np.int64)
    d["spec_residues"] = np.zeros(0, np.int64)
    d["spec_only_k0"] = np.zeros(0, np.int64)
    d["spec_only_irrep"] = np.zeros(0, np.int64)
    towers = [L["total_spin"] for L in levels if "total_spin" in L]
    two_S = int(round(2 * towers[0])) if towers else -1   # spec_scalars[5]: 2S of a total_spin request
    d["spec_scalars"] = np.array([-1, 0, 1, 0, 0, two_S, result.n_sites], np.int64)
    total = int(sum(r["dim"] for r in result.blocks))
    d["result_scalars"] = np.array([total, 0, int(bool(result.converged)), 0, 0, 0,
                                    int(result.timings.get("pruned_blocks", 0))], np.int64)
    return d


def _fail(what: str):
    raise InvalidRequest("load_eigs: " + what)


def _get(d, key, dtype):
    if key not in d:
        _fail(f"the file has no array {key}")
    try:
        return np.asarray(d[key]).astype(dtype, casting="same_kind" if np.dtype(dtype).kind == "c" else "unsafe").reshape(-1)
    except (TypeError, ValueError):
        _fail(f"bad array {key}")


def _check_offsets(off, n, size, name):
    if off.size != n + 1:
        _fail(f"{name} has the wrong length")
    if np.any(off < 0) or np.any(off > size) or np.any(np.diff(off) < 0):
        _fail(f"{name} is not a list of offsets into its array")


def from_arrays(d: Dict[str, np.ndarray]):
    """(levels, bases, vectors, scalars) of a file's arrays (eigs_from_arrays,
    python/qed/_bindings/sectors.cpp:163-330), every index checked."""
    energy = _get(d, "level_energy", np.float64)
    mult = _get(d, "level_multiplicity", np.int64)
    vector = _get(d, "level_vector", np.int64)
    tag = _get(d, "level_tag", np.int64)
    nl = energy.size
    if tag.size != TAG_WIDTH * nl:
        _fail("level_tag has the wrong shape")
    if mult.size != nl or vector.size != nl:
        _fail("the level arrays differ in length")
    tag = tag.reshape(nl, TAG_WIDTH)
    momenta = [None] * nl
    if "level_momentum" in d:
        mchi = _get(d, "level_momentum", np.complex128)
        moff = _get(d, "level_momentum_offset", np.int64)
        _check_offsets(moff, nl, mchi.size, "level_momentum_offset")
        momenta = [mchi[moff[i]:moff[i + 1]] for i in range(nl)]
    nb = int(_get(d, "n_bases", np.int64)[0]) if "n_bases" in d else 0
    if nb < 0:
        _fail("n_bases is negative")
    bases = []
    for b in range(nb):
        p = f"basis{b}_"
        shape = _get(d, p + "shape", np.int64)
        if shape.size < 3 or shape[0] < 1 or shape[1] < 1 or shape[1] > 63 or shape[2] < -1 or shape[2] > shape[1]:
            _fail(f"basis {b} has an impossible shape")
        G, N, n_up = int(shape[0]), int(shape[1]), int(shape[2])
        irrep_dim = int(shape[3]) if shape.size > 3 else 1
        if irrep_dim != 1:
            raise InvalidRequest(f"load_eigs: basis {b} is a sector of an irrep of dimension {irrep_dim}; "
                                 "only one-dimensional irreps are supported")
        perms = _get(d, p + "perms", np.int64)
        if perms.size != G * N:
            _fail(f"{p}perms has the wrong shape")
        perms = perms.reshape(G, N)
        for g in range(G):
            if np.any(perms[g] < 0) or np.any(perms[g] >= N) or np.unique(perms[g]).size != N:
                _fail(f"{p}perms holds an entry that is not a permutation of the sites")
        chars = _get(d, p + "characters", np.complex128)
        reps = _get(d, p + "reps", np.uint64)
        inv = _get(d, p + "inv_norms", np.float64)
        flips = _get(d, p + "flip_masks", np.uint64) if (p + "flip_masks") in d else np.zeros(0, np.uint64)
        if chars.size != G or inv.size != reps.size or (flips.size and flips.size != G):
            _fail(f"basis {b} has arrays of the wrong length")
        if not flips.size:
            flips = np.zeros(G, np.uint64)
        slc = int(_get(d, p + "sublattice", np.uint64)[0]) if (p + "sublattice") in d else 0
        bases.append(LevelBasis(perms=perms.astype(np.int32), chars=chars, flips=flips, n_up=n_up, reps=reps,
                                inv_norms=inv, sublattice=slc))
    vbasis = _get(d, "vector_basis", np.int64) if "vector_basis" in d else np.zeros(0, np.int64)
    voff = _get(d, "vector_offset", np.int64) if "vector_offset" in d else np.zeros(1, np.int64)
    amps = _get(d, "vector_amplitudes", np.complex128) if "vector_amplitudes" in d else np.zeros(0, np.complex128)
    _check_offsets(voff, vbasis.size, amps.size, "vector_offset")
    vectors = []
    for i in range(vbasis.size):
        if vbasis[i] < 0 or vbasis[i] >= nb:
            _fail("vector_basis names a basis that is not in the file")
        v = amps[voff[i]:voff[i + 1]].copy()
        if v.size != bases[vbasis[i]].dim:
            _fail("a vector does not match its basis")
        vectors.append((int(vbasis[i]), v))
    if np.any(vector < -1) or np.any(vector >= len(vectors)):
        _fail("a level names a vector that is not in the file")
    spec = _get(d, "spec_scalars", np.int64) if "spec_scalars" in d else np.zeros(0, np.int64)
    n_sites = int(spec[6]) if spec.size > 6 else (bases[0].N if bases else 0)
    two_S = int(spec[5]) if spec.size > 5 else -1
    levels = []
    for i in range(nl):
        t = tag[i]
        mom = None
        if momenta[i] is not None and momenta[i].size == 1 and n_sites > 0:
            mom = int(round(np.angle(momenta[i][0]) * n_sites / (2 * np.pi))) % n_sites
        levels.append({"energy": float(energy[i]), "n_up": int(t[0]), "momentum": mom,
                       "parity": None if t[5] < 0 else (1 if t[5] == 0 else -1), "dim": int(t[9]),
                       "multiplicity": int(mult[i]), "vector": int(vector[i]), "label": f"n_up={int(t[0])}"})
        if two_S >= 0:
            levels[-1]["total_spin"] = two_S / 2
    rs = _get(d, "result_scalars", np.int64) if "result_scalars" in d else np.zeros(7, np.int64)
    return levels, bases, vectors, n_sites, rs
