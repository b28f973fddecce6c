"""Write the golden sector fixtures from the reference's own header-only hot path.

Runs oracle/_ref/ref_driver (built by `make -C oracle` from the reference checkout's
include/ed/{basis,ops,matvec} and src/ops/{algebra,program}.cpp) on small cases and stores what it
prints as tests/golden/fixtures/<case>.npz: representatives, 1/norms, the sublattice key order,
the row program of H^dagger, index_and_projection of probe states, the merged reduced CSR and one
walk-lane matvec. The fixtures are data; the reference's source stays out of this repository.

    python tests/golden/make_fixtures.py          # needs oracle/_ref/ref_driver
"""

from __future__ import annotations

import math
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
OUT = os.path.join(HERE, "fixtures")

sys.path.insert(0, ROOT)


def root_of_unity(p, L):
    p %= L
    if (4 * p) % L == 0:
        return [1 + 0j, 1j, -1 + 0j, -1j][(4 * p) // L]
    th = 6.283185307179586476925286766559 * float(p) / float(L)
    return complex(math.cos(th), math.sin(th))


def translations(N):
    return [[(i + a) % N for i in range(N)] for a in range(N)]


def reflections(N):
    return [[(a - i) % N for i in range(N)] for a in range(N)]


def heis(bonds, J=1.0, Jz=None):
    Jz = J if Jz is None else Jz
    out = []
    for i, j in bonds:
        out += [("+-", (i, j), 0.5 * J), ("-+", (i, j), 0.5 * J), ("zz", (i, j), Jz)]
    return out


def chiral(tris, chi):
    out = []
    for i, j, k in tris:
        for a, b, c, s in (("x", "y", "z", 1), ("y", "z", "x", 1), ("z", "x", "y", 1), ("x", "z", "y", -1),
                           ("y", "x", "z", -1), ("z", "y", "x", -1)):
            out.append((a + b + c, (i, j, k), s * chi))
    return out


def ring(N, d=1):
    return [(i, (i + d) % N) for i in range(N)]


def case_list():
    cases = {}
    j1j2 = lambda N: heis(ring(N)) + heis(ring(N, 2), 0.5)
    # J1-J2 ring, translations at k = 0, plain order
    cases["j1j2_n12_k0"] = dict(N=12, n_up=6, perms=translations(12), chars=[1.0] * 12, flips=None, mode=0,
                                records=j1j2(12))
    # complex characters
    cases["j1j2_n12_k3"] = dict(N=12, n_up=6, perms=translations(12),
                                chars=[root_of_unity(3 * a, 12) for a in range(12)], flips=None, mode=0,
                                records=j1j2(12))
    cases["j1j2_n12_k5"] = dict(N=12, n_up=6, perms=translations(12),
                                chars=[root_of_unity(5 * a, 12) for a in range(12)], flips=None, mode=0,
                                records=j1j2(12))
    # dihedral, reflection-odd at k = 0
    cases["j1j2_n12_d_k0_odd"] = dict(N=12, n_up=6, perms=translations(12) + reflections(12),
                                      chars=[1.0] * 12 + [-1.0] * 12, flips=None, mode=0, records=j1j2(12))
    # flip-extended translations, flip-odd at k = 6 (pi)
    G = translations(12)
    cases["heis_n12_flip_k6"] = dict(N=12, n_up=6, perms=G + G,
                                     chars=[root_of_unity(6 * a, 12) for a in range(12)]
                                     + [-root_of_unity(6 * a, 12) for a in range(12)],
                                     flips=[0] * 12 + [(1 << 12) - 1] * 12, mode=0, records=heis(ring(12)))
    # sublattice key order forced on (two blocks of 8 sites), dihedral k = 0 even, and k = 3
    cases["j1j2_n16_d_k0_slc"] = dict(N=16, n_up=8, perms=translations(16) + reflections(16),
                                      chars=[1.0] * 32, flips=None, mode=1, records=j1j2(16))
    cases["j1j2_n16_k3_slc"] = dict(N=16, n_up=8, perms=translations(16),
                                    chars=[root_of_unity(3 * a, 16) for a in range(16)], flips=None, mode=1,
                                    records=j1j2(16))
    cases["j1j2_n16_k3_plain"] = dict(N=16, n_up=8, perms=translations(16),
                                      chars=[root_of_unity(3 * a, 16) for a in range(16)], flips=None, mode=0,
                                      records=j1j2(16))
    # XXZ away from half filling, and a time-reversal-odd chirality model (complex H)
    cases["xxz_n10_nup4_k1"] = dict(N=10, n_up=4, perms=translations(10),
                                    chars=[root_of_unity(a, 10) for a in range(10)], flips=None, mode=0,
                                    records=heis(ring(10), 1.0, 0.5))
    tris = [(i, (i + 1) % 10, (i + 2) % 10) for i in range(10)]
    cases["chiral_n10_k2"] = dict(N=10, n_up=5, perms=translations(10),
                                  chars=[root_of_unity(2 * a, 10) for a in range(10)], flips=None, mode=0,
                                  records=j1j2(10) + chiral(tris, 0.25))
    # the full 2^N space with translations (no U(1))
    cases["heis_n8_full_k1"] = dict(N=8, n_up=-1, perms=translations(8),
                                    chars=[root_of_unity(a, 8) for a in range(8)], flips=None, mode=0,
                                    records=heis(ring(8)))
    return cases


def write_case_file(path, c, probes):
    N, G = c["N"], len(c["perms"])
    flips = c["flips"] or [0] * G
    with open(path, "w") as f:
        f.write(f"{N} {c['n_up']} {G} {c['mode']} 1\n")
        for p in c["perms"]:
            f.write(" ".join(map(str, p)) + "\n")
        for m in flips:
            f.write(f"{m}\n")
        for ch in c["chars"]:
            ch = complex(ch)
            f.write(f"{ch.real!r} {ch.imag!r}\n")
        f.write(f"{len(c['records'])}\n")
        for ops, sites, coeff in c["records"]:
            coeff = complex(coeff)
            f.write(f"{ops} {' '.join(map(str, sites))} {coeff.real!r} {coeff.imag!r}\n")
        f.write(f"{len(probes)}\n")
        for s in probes:
            f.write(f"{int(s)}\n")


def probe_states(c, count=400, seed=5):
    N, n_up = c["N"], c["n_up"]
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        if n_up >= 0:
            bits = rng.permutation(N)[:n_up]
            s = sum(1 << int(b) for b in bits)
        else:
            s = int(rng.integers(0, 1 << N))
        out.append(s)
    return out


def read(d, name, dtype):
    p = os.path.join(d, name)
    return np.fromfile(p, dtype=dtype) if os.path.exists(p) else None


def main():
    if not os.path.exists(DRIVER):
        sys.exit(f"{DRIVER} is not built (make -C oracle)")
    os.makedirs(OUT, exist_ok=True)
    for name, c in case_list().items():
        probes = probe_states(c)
        with tempfile.TemporaryDirectory() as tmp:
            cf = os.path.join(tmp, "case.txt")
            write_case_file(cf, c, probes)
            r = subprocess.run([DRIVER, cf, tmp], capture_output=True, text=True, check=True)
            arrays = dict(
                N=np.int64(c["N"]),
                n_up=np.int64(c["n_up"]),
                perms=np.asarray(c["perms"], np.int32),
                flips=np.asarray(c["flips"] or [0] * len(c["perms"]), np.uint64),
                chars=np.asarray([complex(x) for x in c["chars"]], np.complex128),
                sublattice_mode=np.int64(c["mode"]),
                rec_ops=np.asarray([o for o, _, _ in c["records"]]),
                rec_sites=np.asarray([list(s) + [-1] * (3 - len(s)) for _, s, _ in c["records"]], np.int32),
                rec_coeff=np.asarray([complex(x) for _, _, x in c["records"]], np.complex128),
                probes=np.asarray(probes, np.uint64),
                reps=read(tmp, "reps.u64", np.uint64),
                inv_norms=read(tmp, "inv_norms.f64", np.float64),
                n_orbits=read(tmp, "n_orbits.i64", np.int64),
                program_group_flip=read(tmp, "program_group_flip.u64", np.uint64),
                program_group_setbits=read(tmp, "program_group_setbits.i32", np.int32),
                program_group_vbegin=read(tmp, "program_group_vbegin.u32", np.uint32),
                program_vsub_val=read(tmp, "program_vsub_val.u64", np.uint64),
                program_vsub_tbegin=read(tmp, "program_vsub_tbegin.u32", np.uint32),
                program_term_sign=read(tmp, "program_term_sign.u64", np.uint64),
                program_term_coeff=read(tmp, "program_term_coeff.f64", np.float64),
                probe_index=read(tmp, "probe_index.i64", np.int64),
                probe_proj=read(tmp, "probe_proj.f64", np.float64),
                csr_row_ptr=read(tmp, "csr_row_ptr.u64", np.uint64),
                csr_col=read(tmp, "csr_col.u32", np.uint32),
                csr_val=read(tmp, "csr_val.f64", np.float64),
                csr_meta=read(tmp, "csr_meta.i64", np.int64),
                gather_in=read(tmp, "gather_in.f64", np.float64),
                gather_out=read(tmp, "gather_out.f64", np.float64),
            )
            kb = read(tmp, "key_bit.i32", np.int32)
            if kb is not None:
                arrays["key_bit"] = kb
            np.savez_compressed(os.path.join(OUT, name + ".npz"), **arrays)
            print(name, r.stdout.strip())


if __name__ == "__main__":
    main()
