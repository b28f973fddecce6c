// =============================================================================
// oracle/ref_driver.cpp -- golden-vector writer over the reference's own header-only hot path.
//
// TEST INFRASTRUCTURE ONLY. Built by oracle/Makefile from the reference checkout's headers
// (include/ed/basis, include/ed/ops, include/ed/matvec) and its src/ops/{algebra,program}.cpp,
// into oracle/_ref/ref_driver; never linked into the product. tests/golden/make_fixtures.py runs
// it on small cases and commits what it writes as fixtures.
//
// Input: a case file (text) naming the sector and the operator:
//   N n_up G sublattice_mode(-1 unset, 0 off, 1 on) relaxed(0|1)
//   G lines: the permutation of element g (N ints)
//   G lines: the XOR flip mask of element g
//   G lines: chi(g) re im
//   R: then R record lines  "<ops> <sites...> <re> <im>"   (ops from + - z, one char per site)
//   P: then P probe states (index_and_projection)
// Output (raw little-endian arrays in <outdir>): reps.u64, inv_norms.f64, n_orbits.i64,
// key_bit.i32 (when a sublattice code engages), program_*.{u64,i32,u32,f64}, probe_index.i64,
// probe_proj.f64 (re, im interleaved), csr_row_ptr.u64, csr_col.u32, csr_val.f64 (re, im),
// csr_meta.i64 (nnz, dict size, id width), gather_in.f64, gather_out.f64.
// =============================================================================
#include <ed/basis/compiled_group.h>
#include <ed/basis/orbit_table.h>
#include <ed/basis/rep_sector.h>
#include <ed/basis/sublattice_code.h>
#include <ed/matvec/sector_rows.h>
#include <ed/ops/algebra.h>
#include <ed/ops/program.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

namespace {

template <class T> void dump(const std::string& dir, const std::string& name, const T* p, std::size_t n) {
    const std::string path = dir + "/" + name;
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) {
        std::fprintf(stderr, "cannot write %s\n", path.c_str());
        std::exit(2);
    }
    if (n) std::fwrite(p, sizeof(T), n, f);
    std::fclose(f);
}

template <class T> void dumpv(const std::string& dir, const std::string& name, const std::vector<T>& v) {
    dump(dir, name, v.data(), v.size());
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: ref_driver <case file> <outdir>\n");
        return 2;
    }
    std::ifstream in(argv[1]);
    const std::string out = argv[2];
    int N, n_up, G, mode, relaxed;
    in >> N >> n_up >> G >> mode >> relaxed;
    std::vector<std::vector<int>> perms(G, std::vector<int>(N));
    for (auto& p : perms)
        for (auto& x : p) in >> x;
    std::vector<std::uint64_t> flips(G);
    for (auto& f : flips) in >> f;
    std::vector<std::complex<double>> chi(G);
    for (auto& c : chi) {
        double re, im;
        in >> re >> im;
        c = {re, im};
    }
    // The operator: records summed into canonical terms like ed::ops::masked (src/ops/invariance.cpp:43-60).
    int R;
    in >> R;
    ed::ops::MaskedOperator H(N);
    for (int r = 0; r < R; ++r) {
        std::string ops;
        in >> ops;
        std::vector<int> sites(ops.size());
        for (auto& s : sites) in >> s;
        double re, im;
        in >> re >> im;
        H.add(ed::ops::MaskedOperator::product(N, ops, sites, {re, im}));
    }
    int P;
    in >> P;
    std::vector<std::uint64_t> probes(P);
    for (auto& p : probes) in >> p;

    if (mode >= 0) setenv("ED_SYM_SUBLATTICE", mode ? "1" : "0", 1);
    else unsetenv("ED_SYM_SUBLATTICE");
    ed::symmetry::SublatticeRuleScope rule(relaxed != 0);

    bool any_flip = false;
    for (auto f : flips) any_flip = any_flip || f != 0;
    const auto cg = ed::symmetry::CompiledGroup::from_elements(perms, flips, N);
    ed::symmetry::OrbitTable tab;
    if (n_up >= 0) tab = ed::symmetry::build_orbit_table_fixed_sz_streaming(N, n_up, cg);
    else tab = ed::symmetry::build_orbit_table_full_compiled(N, cg);

    ed::symmetry::RepSectorData rd;
    rd.n_sites = N;
    rd.n_up = n_up;
    rd.group_size = G;
    rd.characters = chi;
    for (const auto& p : perms) rd.perms_flat.insert(rd.perms_flat.end(), p.begin(), p.end());
    if (any_flip) rd.flip_masks = flips;
    rd.slc = tab.slc;
    // The character filter with the reference's norm (orbit_table.h projected_norm_sq_stab) and the
    // keep rule of group_sector.cpp filter_reps.
    for (std::size_t i = 0; i < tab.reps.size(); ++i) {
        const double nsq = ed::symmetry::projected_norm_sq_stab(tab.stabilizer_of(i), chi);
        const double inv = nsq <= 1e-12 ? 0.0 : 1.0 / std::sqrt(nsq);
        if (inv == 0.0) continue;
        rd.reps.push_back(tab.reps[i]);
        rd.inv_norms.push_back(inv);
    }
    rd.build_perm_lut();
    rd.build_buckets();
    const std::int64_t n_orbits = static_cast<std::int64_t>(tab.reps.size());
    dumpv(out, "reps.u64", rd.reps);
    dumpv(out, "inv_norms.f64", rd.inv_norms);
    dump(out, "n_orbits.i64", &n_orbits, 1);
    if (tab.slc) {
        std::vector<std::int32_t> kb(tab.slc->key_bit().begin(), tab.slc->key_bit().end());
        dumpv(out, "key_bit.i32", kb);
    }

    // Program of H^dagger, as Operator::row_program (include/ed/ops/operator.h:132-138).
    const ed::ops::MaskedProgram prog = ed::ops::compile_operator(H.dagger());
    dumpv(out, "program_group_flip.u64", prog.group_flip);
    std::vector<std::int32_t> setbits(prog.group_setbits.begin(), prog.group_setbits.end());
    dumpv(out, "program_group_setbits.i32", setbits);
    dumpv(out, "program_group_vbegin.u32", prog.group_vbegin);
    dumpv(out, "program_vsub_val.u64", prog.vsub_val);
    dumpv(out, "program_vsub_tbegin.u32", prog.vsub_tbegin);
    dumpv(out, "program_term_sign.u64", prog.term_sign);
    std::vector<double> coeff;
    for (const auto& c : prog.term_coeff) {
        coeff.push_back(c.real());
        coeff.push_back(c.imag());
    }
    dumpv(out, "program_term_coeff.f64", coeff);

    const auto pol = rd.make_policy();
    std::vector<std::int64_t> pidx(P);
    std::vector<double> pproj(2 * static_cast<std::size_t>(P));
    for (int i = 0; i < P; ++i) {
        std::complex<double> pr(0.0, 0.0);
        pidx[i] = pol.index_and_projection(probes[i], pr);
        pproj[2 * i] = pr.real();
        pproj[2 * i + 1] = pr.imag();
    }
    dumpv(out, "probe_index.i64", pidx);
    dumpv(out, "probe_proj.f64", pproj);

    const auto view = prog.view();
    const auto csr = ed::matvec::build_sector_csr(view, pol, rd.dim());
    std::vector<std::uint64_t> row_ptr(csr.row_ptr.begin(), csr.row_ptr.end());
    std::vector<std::uint32_t> col(csr.col_idx.begin(), csr.col_idx.end());
    std::vector<double> val;
    for (std::uint64_t e = 0; e < csr.nnz(); ++e) {
        const auto v = csr.value(e);
        val.push_back(v.real());
        val.push_back(v.imag());
    }
    const std::int64_t width = !csr.id8.empty() ? 1 : (!csr.id16.empty() ? 2 : 16);
    const std::int64_t meta[3] = {static_cast<std::int64_t>(csr.nnz()), static_cast<std::int64_t>(csr.dict.size()),
                                  width};
    dumpv(out, "csr_row_ptr.u64", row_ptr);
    dumpv(out, "csr_col.u32", col);
    dumpv(out, "csr_val.f64", val);
    dump(out, "csr_meta.i64", meta, 3);

    // The walk lane on a fixed input (sector_rows.h sector_gather).
    std::mt19937_64 gen(12345);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    std::vector<std::complex<double>> x(rd.dim()), y(rd.dim());
    for (auto& v : x) {
        const double re = u(gen);
        const double im = u(gen);
        v = {re, im};
    }
    ed::matvec::sector_gather(view, pol, rd.dim(), x.data(), y.data());
    dump(out, "gather_in.f64", reinterpret_cast<const double*>(x.data()), 2 * x.size());
    dump(out, "gather_out.f64", reinterpret_cast<const double*>(y.data()), 2 * y.size());
    std::printf("dim %llu orbits %lld nnz %llu dict %zu\n", static_cast<unsigned long long>(rd.dim()),
                static_cast<long long>(n_orbits), static_cast<unsigned long long>(csr.nnz()), csr.dict.size());
    return 0;
}
