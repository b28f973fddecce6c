"""The eigs verb on the HIP backend: the k lowest levels of H over the blocks of a symmetry walk (the
reference's qed.eigs, python/qed/_verbs/eigs.py, driving src/engine/eigs.cpp:360-450 per block).

Per block: want = min(k, dim) levels; blocks of at most two states, or at most the dense floor, are
solved densely (the reference's dense lanes, src/engine/block_solve.cpp:164-262: host pool, or the
device from 1024 states); the others on the device, k = 1 by the lowest-level scan and k > 1 by
Krylov-Schur (block_solve.cpp:263-655). Blocks above 64 states are first estimated and pruned as
the reference's walk prunes them (walk_levels). The levels of every solved block are merged and the
lowest k returned, ascending.

The expect verb (python/qed/_verbs/expect.py) returns, for each of the k lowest levels, the
expectation values <v|O|v> of observables in the level's eigenvector v (the reference sweeps the
observables' programs over the level vectors, rep_matrix_elements, include/ed/ops/program.h:135-160):
here each observable's lambda-projected program between two handles of the level's sector is
applied to v on the device (gather lane) and <v|O v> / <v|v> reduced on the device.

The spectrum verb (python/qed/_verbs/spectrum.py; the DenseBatch walk of src/engine/walk.h) solves
every block densely -- from 1024 states on the device (rocSOLVER zheevd, batched per dimension),
below on the host pool -- and returns all eigenvalues, ascending.
"""

from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import _lib
from . import levels as lv
from .errors import InvalidRequest
from .operator import Operator
from .sector import DEFAULT_CSR_BUDGET, LANE_CSR, LANE_GATHER, DeviceArray, Sector, SectorMatvec, parallel_build
from .symmetry import Symmetry
from .thermal import _handles, release_later, resolve_device, time_reversal_partners, wait_released, walk_blocks

# the automatic dense floor's ceiling (src/engine/internal.h:815)
AUTO_DENSE_CEILING = 8192
# dense blocks from this size are solved on the device (kDeviceDenseMinDim, include/ed/core/device.h)
DEVICE_DENSE_MIN_DIM = 1024
# the dense lanes' largest block
DENSE_MAX_DIM = 32768


def dense_floor(k: int, dense_max_dim: Optional[int]) -> int:
    """The largest block solved densely (lowest_dense_floor, block_solve.cpp:164-197): an explicit
    dense_max_dim as given; automatically 4 max(40 k, 400), at most 8192. (The reference also caps
    it by the host memory a dense solve may take; at 8192 states that is 1.6 GB, always available.)"""
    if dense_max_dim is not None and dense_max_dim >= 0:
        return min(int(dense_max_dim), AUTO_DENSE_CEILING)
    return min(4 * max(40 * int(k), 400), AUTO_DENSE_CEILING)


@dataclass
class EigsResult:
    """The k lowest levels (energies, ascending) and the walk's blocks. `levels`: one dict per
    returned level (energy, n_up, momentum, parity, dim, label, the index of its vector or -1).
    With vectors=True each level's eigenvector is kept in its sector basis: vectors() expands them,
    expect() and correlations() evaluate observables in them, save() writes the result to an .npz
    file that load_eigs() restores."""

    energies: np.ndarray
    blocks: List[dict] = field(default_factory=list)
    converged: bool = True
    timings: Dict[str, float] = field(default_factory=dict)
    k: int = 0
    n_sites: int = 0
    levels: List[dict] = field(default_factory=list)
    _bases: Dict[int, "lv.LevelBasis"] = field(default_factory=dict, repr=False)
    _vectors: List[tuple] = field(default_factory=list, repr=False)   # (basis key, sector vector)
    _cache: Dict = field(default_factory=dict, repr=False, compare=False)

    def _need_vectors(self):
        if not any(L.get("vector", -1) >= 0 for L in self.levels):
            raise InvalidRequest("no vectors: call eigs(..., vectors=True)")

    def vectors(self, basis: str = "full", n_up: Optional[int] = None) -> List[np.ndarray]:
        """Unit eigenvectors of the returned levels, in level order: basis="full" the 2^N
        computational basis (bit i of a state is site i, set = up), basis="sz" the Sz sector n_up
        (its states in ascending order); levels of another n_up have no component there and are
        left out (the reference's EigResult.vectors, python/qed/_verbs/eigs.py:52-78). Vectors of
        one degenerate level come from one block's eigensolver and are orthonormal; vectors of
        different blocks are orthogonal by their quantum numbers. A level of a spin tower
        (total_spin=S) is a whole multiplet: basis="full" gives its members from Sz = S down (as many as
        the k energies still need), basis="sz" the member of that n_up, each lowered from the solved
        Sz = S member by S^-."""
        self._need_vectors()
        if basis not in ("full", "sz"):
            raise InvalidRequest("basis must be 'full' or 'sz'")
        if basis == "sz" and n_up is None:
            raise InvalidRequest("basis='sz' needs n_up")
        out = []
        for L in self.levels:
            if L.get("vector", -1) < 0:
                continue
            key, v = self._vectors[L["vector"]]
            B = self._bases[key]
            if "total_spin" in L and L.get("multiplicity", 1) > 1:
                bottom = self.n_sites - B.n_up          # n_up of the Sz = -S member
                if basis == "full":
                    if len(out) >= self.k:
                        break
                    lowest = max(bottom, B.n_up - (self.k - len(out)) + 1)
                    out.extend(lv.tower_members(B, v, True, lowest, self._cache.setdefault(("tower", key), {})))
                elif bottom <= int(n_up) <= B.n_up:
                    out.append(lv.tower_members(B, v, False, int(n_up),
                                                self._cache.setdefault(("tower", key), {}))[-1])
                continue
            if basis == "sz" and B.n_up != int(n_up):
                continue
            out.extend(lv.expand(B, [v], full=(basis == "full")))
        return out

    def expect(self, ops) -> np.ndarray:
        """<v|O|v> of each level's vector v for every O: a complex array [levels, ops]."""
        self._need_vectors()
        obs = [ops] if isinstance(ops, Operator) else list(ops)
        out = np.zeros((len(self.levels), len(obs)), np.complex128)
        for i, L in enumerate(self.levels):
            if L.get("vector", -1) < 0:
                out[i] = np.nan
                continue
            key, v = self._vectors[L["vector"]]
            out[i] = lv.expect_in(self._bases[key], v, obs, self._cache.setdefault(("expect", key), {}))
        return out

    def correlations(self, A, B=None) -> np.ndarray:
        """<A_a^dag B_b> in each level's vector for every pair (B=None: B = A): a complex array
        [levels, len(A), len(B)] (the reference's EigResult.correlations,
        python/qed/_verbs/eigs.py:91-98, per returned vector)."""
        self._need_vectors()
        As = [A] if isinstance(A, Operator) else list(A)
        Bs = As if B is None else ([B] if isinstance(B, Operator) else list(B))
        out = np.zeros((len(self.levels), len(As), len(Bs)), np.complex128)
        for i, L in enumerate(self.levels):
            if L.get("vector", -1) < 0:
                out[i] = np.nan
                continue
            key, v = self._vectors[L["vector"]]
            out[i] = lv.correlations_in(self._bases[key], v, As, Bs, self._cache.setdefault(("corr", key), {}))
        return out

    def save(self, path) -> None:
        """Write the result to an .npz file (format_version 2: the levels, each basis once, the
        vectors in their sector basis); load_eigs() restores it and vectors(), expect() and
        correlations() then work without H."""
        np.savez_compressed(path, **lv.to_arrays(self))


def load_eigs(path) -> EigsResult:
    """An EigsResult written by EigsResult.save (python/qed/_verbs/eigs.py:191-215): format 1 (the
    old spin convention) and files that are not EigResult files are refused; the bases are imported
    onto the device when a vector is first used."""
    with np.load(path, allow_pickle=False) as f:
        d = {key: f[key] for key in f.files}
    version = int(d["format_version"]) if "format_version" in d else 0
    if version == 1:
        raise InvalidRequest(f"{path}: an EigResult file of format 1, written when a set bit meant spin down; "
                             "its states and n_up labels mean the opposite now -- recompute it")
    if version != lv.FORMAT_VERSION:
        raise InvalidRequest(f"{path}: not an EigResult file (format_version 2)")
    levels, bases, vectors, n_sites, rs = lv.from_arrays(d)
    energies = np.asarray(d["energies"], np.float64).reshape(-1) if "energies" in d else np.zeros(0)
    k = int(d["k"]) if "k" in d else len(energies)
    return EigsResult(energies=energies, blocks=[], converged=bool(rs[2]) if rs.size > 2 else True,
                      timings={"pruned_blocks": int(rs[6]) if rs.size > 6 else 0}, k=k, n_sites=n_sites,
                      levels=levels, _bases={i: b for i, b in enumerate(bases)}, _vectors=vectors)


PRUNE_MIN_DIM = 64   # blocks above this are estimated before they are solved (eigs.cpp:129-132)


def _solve_block(m, dim: int, want: int, floor: int, max_iter: int):
    """The `want` lowest levels of one block: densely (host or device lane) at or below the dense
    floor, else the device Krylov lanes. (values, converged, lane)"""
    if dim <= 2 or dim <= floor:
        vals = np.zeros(dim)
        _lib.call("qedh_block_spectrum", _handles([m]), 1, C.c_int64(-1), _lib.ptr(vals, C.c_double))
        return vals[:want].copy(), True, "device-dense" if dim >= DEVICE_DENSE_MIN_DIM else "host-dense"
    v, conv, _ = m.lowest(want, max_iter)
    return np.asarray(v), conv, "device-krylov"


def walk_order(est, cand, partners) -> list:
    """Positions q of cand in the order a pruned walk takes them: increasing estimate est[q], ties by
    block index, each mirror of a time-reversal pair (partners {mirror: block}) right after its
    partner whatever the two blocks' indices -- the mirror shares the partner's estimate and then its
    decision (solved or pruned), so the walk does not depend on the order the momenta were given in."""
    def key(q):
        i = cand[q]
        return (est[q], partners.get(i, i), i in partners)
    return sorted(range(len(cand)), key=key)


def walk_levels(blocks, k: int, *, floor: int, prune: bool = True, prune_margin: float = 0.02, window: float = 0.0,
                norm_bound: float = 1.0, max_iter: int = 0, partners=None):
    """The levels of the blocks of a walk as the reference's eigs walk finds them
    (src/engine/eigs.cpp:360-560): blocks of at most max(min(dense floor, 64), 2) states are solved
    outright; every larger block first gets its pruning estimate (40 Lanczos steps, theta_1 - |r_1|,
    on host threads with streams of their own), then the blocks are taken in order of increasing
    estimate and solved only when the estimate lies within prune_margin (relative to |E_k|, floored at
    5% of s_H) or `window` of the k-th level found so far. Returns {block index: (levels, converged,
    lane)} for the blocks solved and the number pruned. partners {mirror: block} (time-reversal pairs
    of a real H, indices into blocks): a mirror is neither estimated nor solved; it takes its
    partner's estimate and levels, or is pruned with it (walk_order)."""
    partners = {} if partners is None else partners
    levels = {}
    small = max(min(floor, PRUNE_MIN_DIM), 2)
    outright = [i for i, (_, s, _) in enumerate(blocks) if (not prune or s.dim <= small) and i not in partners]
    cand = [i for i, (_, s, _) in enumerate(blocks) if prune and s.dim > small and i not in partners]
    dense = [i for i in outright if blocks[i][1].dim <= 2 or blocks[i][1].dim <= floor]
    if dense:
        mats = [blocks[i][2] for i in dense]
        dims = [mm.dim for mm in mats]
        vals = np.zeros(int(sum(dims)))
        _lib.call("qedh_block_spectrum", _handles(mats), len(mats), C.c_int64(-1), _lib.ptr(vals, C.c_double))
        at = 0
        for i, d in zip(dense, dims):
            levels[i] = (vals[at: at + min(k, d)].copy(), True,
                         "device-dense" if d >= DEVICE_DENSE_MIN_DIM else "host-dense")
            at += d
    for i in outright:
        if i not in levels:
            blk, sec, m = blocks[i]
            levels[i] = _solve_block(m, sec.dim, min(k, sec.dim), floor, max_iter)
    for j, p in partners.items():      # mirrors of blocks solved outright
        if p in levels:
            levels[j] = levels[p]
    est = parallel_build(lambda i: blocks[i][2].estimate(), cand)
    for j, p in partners.items():      # mirrors of candidates share their estimate
        if p in cand:
            cand.append(j)
            est.append(est[cand.index(p)])
    margin_floor = 0.05 * (norm_bound if norm_bound > 0 else 1.0)
    pruned = 0
    for q in walk_order(est, cand, partners):
        i = cand[q]
        if i in partners:    # one spectrum with its partner, which was just taken or pruned
            if partners[i] in levels:
                levels[i] = levels[partners[i]]
            else:
                pruned += 1
            continue
        found = np.sort(np.concatenate([v for v, _, _ in levels.values()])) if levels else np.zeros(0)
        kth = found[k - 1] if found.size >= k else math.inf
        if est[q] > kth + max(prune_margin * max(margin_floor, abs(kth)), window):
            pruned += 1
            continue
        blk, sec, m = blocks[i]
        levels[i] = _solve_block(m, sec.dim, min(k, sec.dim), floor, max_iter)
    return levels, pruned


def _level_vectors(blocks, rows, max_iter: int):
    """Eigenvectors of the selected levels: per block, as many Krylov-Schur eigenpairs on the device
    as the block has levels among the rows (ascending, so the t-th row of a block takes its t-th
    vector). {block: [sector vectors]}, {block: LevelBasis}"""
    from collections import Counter

    got, bases = {}, {}
    for i, c in sorted(Counter(i for _, i in rows).items()):
        blk, sec, m = blocks[i]
        buf = DeviceArray(c * sec.dim)
        vals = np.zeros(c)
        n, conv = C.c_int(), C.c_int()
        _lib.call("qedh_block_eigenpairs", m.h, int(c), C.c_int64(int(max_iter)), _lib.ptr(vals, C.c_double),
                  buf.ptr, C.byref(n), C.byref(conv))
        if n.value < c:
            raise RuntimeError(f"eigs: block {blk.label} gave {n.value} eigenpairs of the {c} asked")
        host = buf.to_host()
        got[i] = [host[j * sec.dim:(j + 1) * sec.dim].copy() for j in range(c)]
        bases[i] = lv.LevelBasis.of(blk.group, sec)
    return got, bases


def spin_squared(N: int) -> Operator:
    """S^2 = 3N/4 + sum_{i<j} (2 S^z_i S^z_j + S^+_i S^-_j + S^-_i S^+_j)."""
    P = Operator(N)
    P.add_product("I", (0,), 0.75 * N)
    for i in range(N):
        for j in range(i + 1, N):
            P.add_product("zz", (i, j), 2.0)
            P.add_product("+-", (i, j), 1.0)
            P.add_product("-+", (i, j), 1.0)
    return P


def two_spin(total_spin, N: int) -> int:
    """2S of a total_spin request: "min" is 0 for even N and 1 for odd N."""
    if isinstance(total_spin, str):
        if total_spin != "min":
            raise InvalidRequest(f"total_spin must be a number or 'min', got {total_spin!r}")
        return N % 2
    two_S = int(round(2.0 * float(total_spin)))
    if abs(2.0 * float(total_spin) - two_S) > 1e-12 or two_S < 0 or two_S > N or (N + two_S) % 2:
        raise InvalidRequest(f"total_spin {total_spin} is not a spin of {N} spins 1/2")
    return two_S


def spin_tower_operator(H: Operator, two_S: int, mu: Optional[float] = None) -> Operator:
    """H + mu (S^2 - S(S+1)) (the reference's penalty operator, src/engine/tower.cpp). H commutes with
    S^2, so every level of spin S keeps its energy; in the Sz = S sector every state has spin >= S and
    a level of spin S' > S rises by mu (S'(S'+1) - S(S+1)) >= 2 (S + 1) mu. The default mu =
    (s_H + 1) / (S + 1) lifts every such level above every level of spin S."""
    N = H.num_sites
    S = 0.5 * two_S
    if mu is None:
        mu = (H.norm_bound() + 1.0) / (S + 1.0)
    lam = S * (S + 1.0)
    out = Operator(N, products=list(H.products))
    for ops, sites, c in spin_squared(N).products:
        out.add_product(ops, sites, mu * c)
    out.add_product("I", (0,), -mu * lam)
    return out


def _resident_bytes(N: int, sym: Symmetry) -> float:
    """A resident walk's device footprint, roughly: per block its rank directory (16 bytes per 64
    states of its Sz subspace) and per state about 48 bytes of representatives, norms and CSR (the blocks'
    states add up to the selected Sz subspaces, or a share of them under momentum blocks)."""
    from math import comb

    total = 0.0
    for b in sym.blocks(N):
        sub = comb(N, b.n_up) if b.n_up >= 0 else 2 ** N
        total += 0.25 * sub + 48.0 * sub * b.weight / max(1, b.group.G)
    return total


def _build_blocks(H: Operator, blks, csr_budget: int, tune: bool):
    """(block, sector, operator) of each block, built on the build threads."""
    secs = parallel_build(lambda b: Sector(b.group, b.n_up), blks)
    mats = parallel_build(lambda s: SectorMatvec(H, s, lane=LANE_CSR, csr_budget=csr_budget, tune=tune), secs)
    return [(b, s, m) for b, s, m in zip(blks, secs, mats)]


def _stream_levels(H: Operator, k: int, sym: Symmetry, *, floor: int, prune: bool, prune_margin: float, window: float,
                   max_iter: int, csr_budget: int):
    """walk_levels without the resident walk: each build thread takes one block at a time -- its
    sector, its operator and (above the outright size) its pruning estimate -- and frees it; the
    blocks solved outright are then built again together and solved as walk_levels solves them, and
    the candidates, by increasing estimate, are built again one by one when not pruned. Returns
    ([(block, dim)], {index: (levels, converged, lane)}, pruned, {}, build seconds, {the candidates'
    rebuild and solve seconds})."""
    t0 = time.perf_counter()
    blks = sym.blocks(H.num_sites)
    H._handle()
    wait_released()
    small = max(min(floor, PRUNE_MIN_DIM), 2)
    # A real H maps the momentum-m block onto the complex conjugate of the momentum-(N - m) block
    # (time reversal: same representatives and norms, conjugate characters), so the two have one
    # spectrum; the second of each such pair is neither built nor solved: it takes its partner's
    # dimension, estimate and levels.
    partner = time_reversal_partners(H, blks)

    def first(b):
        sec = Sector(b.group, b.n_up)
        if sec.dim == 0 or sec.dim <= small:
            return sec.dim, None
        if not prune:        # every larger block is solved, one at a time
            return sec.dim, -math.inf
        return sec.dim, SectorMatvec(H, sec, lane=LANE_CSR, csr_budget=csr_budget, tune=False).estimate()

    primary = [i for i in range(len(blks)) if i not in partner]
    scan = dict(zip(primary, parallel_build(first, [blks[i] for i in primary])))
    for i, p in partner.items():
        scan[i] = scan[p]
    scan = [scan[i] for i in range(len(blks))]
    keep = [i for i, (d, _) in enumerate(scan) if d > 0]
    jmap = {i: j for j, i in enumerate(keep)}
    mirror = {jmap[i]: jmap[p] for i, p in partner.items() if i in jmap}
    if not keep:
        raise InvalidRequest("eigs: the symmetry selection holds no state")
    meta = [(blks[i], scan[i][0]) for i in keep]
    est = {j: scan[i][1] for j, i in enumerate(keep) if scan[i][1] is not None}
    t_build = time.perf_counter() - t0
    out = [j for j in range(len(meta)) if j not in est and j not in mirror]
    levels = {}
    split = {"rebuild_s": 0.0, "rebuild_sector_s": 0.0, "rebuild_csr_s": 0.0, "candidate_solve_s": 0.0}
    if out:
        built = _build_blocks(H, [meta[j][0] for j in out], csr_budget, tune=False)
        lv_out, _ = walk_levels(built, k, floor=floor, prune=False, max_iter=max_iter)
        levels = {out[q]: v for q, v in lv_out.items()}
        del built
    for j, p in mirror.items():
        if j not in est:
            levels[j] = levels[p]
    margin_floor = 0.05 * (H.norm_bound() or 1.0)
    pruned = 0
    ests = sorted(est)
    for q in walk_order([est[j] for j in ests], ests, mirror):
        j = ests[q]
        if j in mirror:      # one spectrum with its partner, which was just taken or pruned
            if mirror[j] in levels:
                levels[j] = levels[mirror[j]]
            else:
                pruned += 1
            continue
        found = np.sort(np.concatenate([v for v, _, _ in levels.values()])) if levels else np.zeros(0)
        kth = found[k - 1] if found.size >= k else math.inf
        if prune and est[j] > kth + max(prune_margin * max(margin_floor, abs(kth)), window):
            pruned += 1
            continue
        blk, dim = meta[j]
        tb = time.perf_counter()
        sec = Sector(blk.group, blk.n_up)
        tm = time.perf_counter()
        m = SectorMatvec(H, sec, lane=LANE_CSR, csr_budget=csr_budget, tune=True)
        ts = time.perf_counter()
        levels[j] = _solve_block(m, dim, min(k, dim), floor, max_iter)
        split["rebuild_s"] += ts - tb
        split["rebuild_sector_s"] += tm - tb
        split["rebuild_csr_s"] += float(m.info.get("build_seconds", 0.0))
        split["candidate_solve_s"] += time.perf_counter() - ts
        del sec, m
    return meta, levels, pruned, {}, t_build, split


def _tower(H: Operator, k: int, sym: Symmetry, two_S: int, **kw) -> EigsResult:
    """The k lowest energies of total spin S (each multiplet counted 2S + 1 times) through the
    penalty operator in the Sz = S sector. A small penalty keeps the operator's spectral width, and
    so the Krylov lanes' convergence, close to H's; it is raised until the levels it returns are
    provably of spin S: every level of a higher spin lies at or above e_min + 2 (S + 1) mu (e_min the
    lowest level of H in the sector), so levels below that are of spin S. At mu = (s_H + 1) / (S + 1)
    the bound exceeds every level of spin S and the answer is final."""
    N = H.num_sites
    S = 0.5 * two_S
    n_up = (N + two_S) // 2
    if sym.sz == "off" or (sym.sz != "all" and [int(x) for x in np.atleast_1d(sym.sz)] != [n_up]):
        raise InvalidRequest(f"eigs: total_spin {S} is solved in the Sz sector n_up = {n_up}; "
                             f"the symmetry asks for sz={sym.sz!r}")
    sym = Symmetry(sz=[n_up], momenta=sym.momenta, translations=sym.translations, reflection=sym.reflection)
    mult = two_S + 1
    k_lv = -(-k // mult)
    t0 = time.perf_counter()
    base = {key: kw[key] for key in ("device", "dense_max_dim", "max_iter", "prune", "prune_margin", "stream",
                                     "csr_budget")}
    e_min = float(eigs(H, 1, sym=sym, **base).energies[0])
    s_H = H.norm_bound()
    mu_safe = (s_H + 1.0) / (S + 1.0)
    mu = mu_safe / 64.0
    while True:
        r = eigs(spin_tower_operator(H, two_S, mu), k_lv, sym=sym, **kw)
        cut = e_min + 2.0 * (S + 1.0) * mu - 1e-9 * max(1.0, s_H)
        good = [L for L in r.levels if L["energy"] < cut]
        if len(good) == len(r.levels) == k_lv or mu >= mu_safe:
            break
        mu = min(mu_safe, 8.0 * mu)
    for L in good:
        L["total_spin"] = S
        L["multiplicity"] = mult
    r.levels = good
    r.energies = np.repeat(np.array([L["energy"] for L in good]), mult)[:k]
    r.k = k
    r.n_sites = N
    r.timings.update({"total_s": time.perf_counter() - t0, "penalty": mu})
    return r


def eigs(H: Operator, k: int = 4, *, sym: Optional[Symmetry] = None, device: str = "gpu",
         dense_max_dim: Optional[int] = None, max_iter: int = 0, prune: bool = True, prune_margin: float = 0.02,
         window: float = 0.0, vectors: bool = False, total_spin=None, stream: Optional[bool] = None,
         csr_budget: int = DEFAULT_CSR_BUDGET) -> EigsResult:
    """The k lowest levels of H over the blocks of `sym` (default: every (Sz, k) block of a ring).
    prune (the reference's default): blocks whose 40-step estimate lies far above the k-th level
    found so far are not solved. vectors: also keep each returned level's eigenvector (Krylov-Schur
    eigenpairs of its block on the device, in the block's sector basis). total_spin (a spin S or
    "min", H SU(2) symmetric): the k lowest energies of total spin S, each multiplet repeated 2S + 1
    times as the reference counts it, solved once in the Sz = S sector of `sym`'s spatial blocks
    under the penalty operator (spin_tower_operator); a level's vector is its Sz = S member.
    stream: hold at most one block per build thread on the device (walks whose blocks do not fit the
    device together); the estimates, the pruning and the levels are those of the resident walk.
    None: stream when the resident walk's estimated footprint exceeds half the free device memory."""
    resolve_device(device, "eigs")
    if int(k) < 1:
        raise InvalidRequest("eigs: k must be >= 1")
    k = int(k)
    sym = sym if sym is not None else Symmetry()
    if total_spin is not None:
        return _tower(H, k, sym, two_spin(total_spin, H.num_sites), device=device, dense_max_dim=dense_max_dim,
                      max_iter=max_iter, prune=prune, prune_margin=prune_margin, window=window, vectors=vectors,
                      stream=stream, csr_budget=csr_budget)
    t0 = time.perf_counter()
    floor = dense_floor(k, dense_max_dim)
    if stream is None:
        stream = _resident_bytes(H.num_sites, sym) > 0.5 * _lib.device_memory()[0]
    if stream:
        meta, levels, pruned, blocks, t_build, split = _stream_levels(H, k, sym, floor=floor, prune=prune,
                                                                      prune_margin=prune_margin, window=window,
                                                                      max_iter=max_iter, csr_budget=csr_budget)
        t1 = t0 + t_build
    else:
        # operators built on first use: time-reversal mirrors and pruned blocks' solves need none
        blocks = walk_blocks(H, sym, csr_budget, tune=False, lazy=True)
        if not blocks:
            raise InvalidRequest("eigs: the symmetry selection holds no state")
        meta = [(blk, sec.dim) for blk, sec, _ in blocks]
        partners = time_reversal_partners(H, [b for b, _, _ in blocks])
        t_build = time.perf_counter() - t0
        t1 = time.perf_counter()
        levels, pruned = walk_levels(blocks, k, floor=floor, prune=prune, prune_margin=prune_margin, window=window,
                                     norm_bound=H.norm_bound(), max_iter=max_iter, partners=partners)
        split = {}
    t_solve = time.perf_counter() - t1
    recs, converged = [], True
    for i, (blk, dim) in enumerate(meta):
        if i not in levels:
            recs.append({"label": blk.label, "dim": dim, "lane": "pruned", "levels": np.zeros(0), "converged": True})
            continue
        v, conv, lane = levels[i]
        converged = converged and conv
        recs.append({"label": blk.label, "dim": dim, "lane": lane, "levels": v, "converged": conv})
    rows = sorted((float(e), i) for i, (v, _, _) in levels.items() for e in v)[:k]
    energies = np.array([e for e, _ in rows])
    t2 = time.perf_counter()
    if vectors and stream:   # the blocks of the returned levels, built again
        need = sorted({i for _, i in rows})
        blocks = dict(zip(need, _build_blocks(H, [meta[i][0] for i in need], csr_budget, tune=True)))
    got, bases = _level_vectors(blocks, rows, max_iter) if vectors else ({}, {})
    lvls, vecs, used = [], [], {}
    for e, i in rows:
        blk, dim = meta[i]
        L = {"energy": e, "n_up": blk.n_up, "momentum": blk.momentum, "parity": blk.parity, "dim": dim,
             "label": blk.label, "vector": -1}
        if vectors:
            t = used.get(i, 0)
            used[i] = t + 1
            L["vector"] = len(vecs)
            vecs.append((i, got[i][t]))
        lvls.append(L)
    if isinstance(blocks, list):
        release_later(blocks)      # the levels' bases keep the sectors they need
    return EigsResult(energies=energies, blocks=recs, converged=converged,
                      timings={"build_s": t_build, "solve_s": t_solve, "pruned_blocks": pruned,
                               "vectors_s": time.perf_counter() - t2 if vectors else 0.0,
                               "total_s": time.perf_counter() - t0, **split},
                      k=k, n_sites=H.num_sites, levels=lvls, _bases=bases, _vectors=vecs)


@dataclass
class SpectrumResult:
    eigenvalues: np.ndarray
    blocks: List[dict] = field(default_factory=list)
    timings: Dict[str, float] = field(default_factory=dict)


def spectrum(H: Operator, *, sym: Optional[Symmetry] = None, device: str = "gpu",
             device_min_dim: Optional[int] = None, csr_budget: int = DEFAULT_CSR_BUDGET) -> SpectrumResult:
    """Every eigenvalue of H over the blocks of `sym`, ascending. device_min_dim: blocks from this
    many states are solved on the device (default 1024; 0 sends every block there). With a real H a
    time-reversal pair of momentum blocks is solved once and its eigenvalues counted twice."""
    resolve_device(device, "spectrum")
    sym = sym if sym is not None else Symmetry()
    t0 = time.perf_counter()
    blocks = walk_blocks(H, sym, csr_budget, tune=False, fold=True)
    if not blocks:
        raise InvalidRequest("spectrum: the symmetry selection holds no state")
    big = [s.dim for _, s, _ in blocks if s.dim > DENSE_MAX_DIM]
    if big:
        raise InvalidRequest(f"spectrum: a block of {big[0]} states is above the dense lanes' {DENSE_MAX_DIM}")
    t_build = time.perf_counter() - t0
    mats = [m for _, _, m in blocks]
    dims = [m.dim for m in mats]
    vals = np.zeros(int(sum(dims)))
    dmin = -1 if device_min_dim is None else int(device_min_dim)
    t1 = time.perf_counter()
    _lib.call("qedh_block_spectrum", _handles(mats), len(mats), C.c_int64(dmin), _lib.ptr(vals, C.c_double))
    t_solve = time.perf_counter() - t1
    recs, at, every = [], 0, []
    lim = DEVICE_DENSE_MIN_DIM if dmin < 0 else dmin
    for (blk, sec, _), d in zip(blocks, dims):
        recs.append({"label": blk.label, "dim": d, "weight": blk.weight,
                     "lane": "device-dense" if (d >= lim or d > 8192) else "host-dense",
                     "eigenvalues": vals[at: at + d].copy()})
        every.extend([vals[at: at + d]] * blk.weight)
        at += d
    release_later(blocks)
    return SpectrumResult(eigenvalues=np.sort(np.concatenate(every)), blocks=recs,
                          timings={"build_s": t_build, "solve_s": t_solve, "total_s": time.perf_counter() - t0})


@dataclass
class ExpectResult:
    energies: np.ndarray          # the k lowest levels, ascending
    values: np.ndarray            # [level, observable]: <v|O|v> / <v|v> (real part)
    labels: List[str] = field(default_factory=list)   # the block of each level
    timings: Dict[str, float] = field(default_factory=dict)


def _vdot(x_ptr, y_ptr, n: int) -> complex:
    out = np.zeros(2)
    _lib.call("qedh_vdot", x_ptr, y_ptr, C.c_int64(int(n)), 1, _lib.ptr(out, C.c_double))
    return complex(out[0], out[1])


def expect(H: Operator, observables, k: int = 4, *, sym: Optional[Symmetry] = None, device: str = "gpu",
           csr_budget: int = DEFAULT_CSR_BUDGET) -> ExpectResult:
    """For each of the k lowest levels of H over the blocks of `sym`: its energy and <v|O|v> of every
    observable O (an Operator or a list) in its eigenvector v (Krylov-Schur eigenpairs per block on
    the device). A level degenerate within its block has no unique vector; the values are then those
    of the vector the solver returns."""
    from .dynamics import _DevView, _lib_alloc

    resolve_device(device, "expect")
    obs = [observables] if isinstance(observables, Operator) else list(observables)
    if not obs or int(k) < 1:
        raise InvalidRequest("expect: at least one observable and k >= 1")
    k = int(k)
    sym = sym if sym is not None else Symmetry()
    t0 = time.perf_counter()
    blocks = walk_blocks(H, sym, csr_budget)
    if not blocks:
        raise InvalidRequest("expect: the symmetry selection holds no state")
    rows, bufs = [], {}
    for i, (blk, sec, m) in enumerate(blocks):
        want = min(k, sec.dim)
        buf = _lib_alloc(want * sec.dim * 16)
        vals = np.zeros(want)
        n, conv = C.c_int(), C.c_int()
        _lib.call("qedh_block_eigenpairs", m.h, int(want), C.c_int64(0), _lib.ptr(vals, C.c_double), buf.ptr,
                  C.byref(n), C.byref(conv))
        bufs[i] = buf
        rows.extend((float(vals[j]), i, j) for j in range(n.value))
    rows.sort(key=lambda r: r[0])
    rows = rows[:k]
    t_solve = time.perf_counter() - t0
    ops, twins = {}, {}
    values = np.zeros((len(rows), len(obs)))
    for li, (e, i, j) in enumerate(rows):
        blk, sec, _ = blocks[i]
        if i not in twins:   # the observables are not group invariant: their lambda-projected programs
            twins[i] = Sector(blk.group, blk.n_up)   # between two handles of the sector
        v = _DevView(bufs[i].ptr.value, j * sec.dim * 16)
        y = _lib_alloc(sec.dim * 16)
        norm2 = _vdot(v.ptr, v.ptr, sec.dim).real
        for o, O in enumerate(obs):
            if (i, o) not in ops:
                ops[(i, o)] = SectorMatvec(O, twins[i], sec, lane=LANE_GATHER)
            ops[(i, o)].apply(v, y, complex_vectors=True)
            values[li, o] = (_vdot(v.ptr, y.ptr, sec.dim) / norm2).real
    return ExpectResult(energies=np.array([r[0] for r in rows]), values=values,
                        labels=[blocks[i][0].label for _, i, _ in rows],
                        timings={"solve_s": t_solve, "total_s": time.perf_counter() - t0})
