// The C-ABI (include/qedh.h): opaque handles over the host engine, exceptions to status codes.
#include "../../../include/qedh.h"

#include <algorithm>
#include <cstring>
#include <memory>
#include <string>

#include "algebra.h"
#include "eigsolve.h"
#include "sector.h"
#include "sector_operator.h"
#include "thermo.h"

struct qedh_operator_s {
    qsa::SpinOperator op;
};
struct qedh_program_s {
    qsa::Program prog;
    std::shared_ptr<qsa::DeviceProgram> dev;   // uploaded on first use
};
struct qedh_sector_s {
    std::shared_ptr<qsa::DeviceSector> sector;
};
struct qedh_matvec_s {
    std::shared_ptr<qsa::SectorOperator> op;
};

namespace {

thread_local std::string g_error;

template <class F> int guarded(F&& f) {
    try {
        g_error.clear();
        f();
        return qsa::kOk;
    } catch (const qsa::Error& e) {
        g_error = e.what();
        return e.code;
    } catch (const std::bad_alloc& e) {
        g_error = std::string("host allocation failed: ") + e.what();
        return qsa::kResourceLimit;
    } catch (const std::invalid_argument& e) {
        g_error = e.what();
        return qsa::kInvalid;
    } catch (const std::exception& e) {
        g_error = e.what();
        return qsa::kRuntime;
    } catch (...) {
        g_error = "unknown error";
        return qsa::kRuntime;
    }
}

void need(const void* p, const char* what) {
    if (!p) throw qsa::Error(qsa::kInvalid, std::string(what) + " is null");
}

std::shared_ptr<qsa::DeviceProgram> device_program(qedh_program p) {
    if (!p->dev) p->dev = std::make_shared<qsa::DeviceProgram>(p->prog);
    return p->dev;
}

}  // namespace

extern "C" {

const char* qedh_last_error(void) { return g_error.c_str(); }

int qedh_device_count(int* count) {
    return guarded([&] {
        need(count, "count");
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess) {
            (void)hipGetLastError();
            n = 0;
        }
        *count = n;
    });
}

int qedh_device_memory(uint64_t* free_bytes, uint64_t* total_bytes) {
    return guarded([&] {
        need(free_bytes, "free_bytes");
        need(total_bytes, "total_bytes");
        qsa::require_device();
        size_t f = 0, t = 0;
        QSA_HIP(hipMemGetInfo(&f, &t));
        *free_bytes = f;
        *total_bytes = t;
    });
}

int qedh_malloc(void** ptr, size_t bytes) {
    return guarded([&] {
        need(ptr, "ptr");
        qsa::require_device();
        *ptr = nullptr;
        if (bytes) QSA_HIP(hipMalloc(ptr, bytes));
    });
}
int qedh_free(void* ptr) {
    return guarded([&] {
        if (ptr) QSA_HIP(hipFree(ptr));
    });
}
int qedh_memcpy_htod(void* dst, const void* src, size_t bytes) {
    return guarded([&] {
        qsa::copy_to_device(dst, src, bytes);
    });
}
int qedh_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
    return guarded([&] {
        qsa::copy_to_host(dst, src, bytes);
    });
}
int qedh_memset(void* dst, int value, size_t bytes) {
    return guarded([&] {
        if (!bytes) return;
        QSA_HIP(hipMemsetAsync(dst, value, bytes, qsa::engine_stream()));
        QSA_HIP(hipStreamSynchronize(qsa::engine_stream()));
    });
}
int qedh_synchronize(void) {
    return guarded([&] {
        QSA_HIP(hipStreamSynchronize(qsa::engine_stream()));
        QSA_HIP(hipDeviceSynchronize());
    });
}
int qedh_thread_stream(int own) {
    return guarded([&] {
        if (!own) {
            qsa::detail::tls_stream() = nullptr;
            return;
        }
        qsa::require_device();
        // taken from the engine's pool on the thread's first request and handed back when the
        // thread ends, with no HIP call at thread exit (a profiler's per-thread state may already
        // be gone then)
        struct Owned {
            hipStream_t s = nullptr;
            ~Owned() { qsa::pooled_stream_return(s); }
        };
        thread_local Owned owned;
        if (!owned.s) owned.s = qsa::pooled_stream_acquire();
        qsa::detail::tls_stream() = owned.s;
    });
}

int qedh_operator_create(int n_sites, qedh_operator* out) {
    return guarded([&] {
        need(out, "out");
        *out = new qedh_operator_s{qsa::SpinOperator(n_sites)};
    });
}
int qedh_operator_add_product(qedh_operator op, const char* ops, const int* sites, double re, double im) {
    return guarded([&] {
        need(op, "operator");
        need(ops, "ops");
        const std::string s(ops);
        std::vector<int> st(sites, sites + s.size());
        op->op.add(qsa::SpinOperator::product(op->op.n_sites(), s, st, qsa::cplx(re, im)));
    });
}
int qedh_operator_info(qedh_operator op, int64_t* n_terms, double* norm_bound, int* is_real) {
    return guarded([&] {
        need(op, "operator");
        if (n_terms) *n_terms = static_cast<int64_t>(op->op.size());
        if (norm_bound) *norm_bound = op->op.l1_norm();
        if (is_real) *is_real = op->op.is_real() ? 1 : 0;
    });
}
int qedh_operator_destroy(qedh_operator op) {
    return guarded([&] { delete op; });
}

int qedh_program_compile(qedh_operator op, int adjoint, qedh_program* out) {
    return guarded([&] {
        need(op, "operator");
        need(out, "out");
        *out = new qedh_program_s{qsa::compile_operator(adjoint ? op->op.dagger() : op->op), nullptr};
    });
}
int qedh_program_compile_projected(qedh_operator op, int adjoint, qedh_sector src, qedh_sector tgt,
                                   qedh_program* out) {
    return guarded([&] {
        need(op, "operator");
        need(src, "src");
        need(tgt, "tgt");
        need(out, "out");
        const auto& a = src->sector->spec();
        const auto& b = tgt->sector->spec();
        if (a.perms != b.perms || a.flips != b.flips) throw qsa::Error(qsa::kInvalid, "program: sectors of different groups");
        const qsa::SpinOperator O = adjoint ? op->op.dagger() : op->op;
        *out = new qedh_program_s{
            qsa::compile_projected(O, a.G, a.perms.data(), a.flips.data(), a.chars.data(), b.chars.data(), a.n_up, b.n_up),
            nullptr};
    });
}
int qedh_program_counts(qedh_program p, int64_t* n_groups, int64_t* n_vsub, int64_t* n_terms) {
    return guarded([&] {
        need(p, "program");
        if (n_groups) *n_groups = static_cast<int64_t>(p->prog.n_groups());
        if (n_vsub) *n_vsub = static_cast<int64_t>(p->prog.vsub_val.size());
        if (n_terms) *n_terms = static_cast<int64_t>(p->prog.n_terms());
    });
}
int qedh_program_export(qedh_program p, uint64_t* group_flip, int32_t* group_setbits, uint32_t* group_vbegin,
                        uint64_t* vsub_val, uint32_t* vsub_tbegin, uint64_t* term_sign, double* term_coeff) {
    return guarded([&] {
        need(p, "program");
        const auto& P = p->prog;
        auto cp = [](auto* dst, const auto& v) {
            if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
        };
        cp(group_flip, P.group_flip);
        cp(group_setbits, P.group_setbits);
        cp(group_vbegin, P.group_vbegin);
        cp(vsub_val, P.vsub_val);
        cp(vsub_tbegin, P.vsub_tbegin);
        cp(term_sign, P.term_sign);
        if (term_coeff)
            for (std::size_t k = 0; k < P.term_coeff.size(); ++k) {
                term_coeff[2 * k] = P.term_coeff[k].real();
                term_coeff[2 * k + 1] = P.term_coeff[k].imag();
            }
    });
}
int qedh_program_import(int64_t n_groups, const uint64_t* group_flip, const int32_t* group_setbits,
                        const uint32_t* group_vbegin, int64_t n_vsub, const uint64_t* vsub_val,
                        const uint32_t* vsub_tbegin, int64_t n_terms, const uint64_t* term_sign,
                        const double* term_coeff, qedh_program* out) {
    return guarded([&] {
        need(out, "out");
        if (n_groups < 0 || n_vsub < 0 || n_terms < 0) throw qsa::Error(qsa::kInvalid, "program: negative count");
        if (n_groups > 0) {
            need(group_flip, "group_flip");
            need(group_setbits, "group_setbits");
        }
        need(group_vbegin, "group_vbegin");
        need(vsub_tbegin, "vsub_tbegin");
        if (n_vsub > 0) need(vsub_val, "vsub_val");
        if (n_terms > 0) {
            need(term_sign, "term_sign");
            need(term_coeff, "term_coeff");
        }
        qsa::Program P;
        P.group_flip.assign(group_flip, group_flip + n_groups);
        P.group_setbits.assign(group_setbits, group_setbits + n_groups);
        P.group_vbegin.assign(group_vbegin, group_vbegin + n_groups + 1);
        P.vsub_val.assign(vsub_val, vsub_val + n_vsub);
        P.vsub_tbegin.assign(vsub_tbegin, vsub_tbegin + n_vsub + 1);
        P.term_sign.assign(term_sign, term_sign + n_terms);
        for (int64_t k = 0; k < n_terms; ++k) P.term_coeff.emplace_back(term_coeff[2 * k], term_coeff[2 * k + 1]);
        // The offsets must be monotone and end at the counts the kernels index with.
        auto monotone = [](const std::vector<std::uint32_t>& v, std::uint64_t end) {
            if (v.empty() || v.front() != 0 || v.back() != end) return false;
            for (std::size_t i = 1; i < v.size(); ++i)
                if (v[i] < v[i - 1]) return false;
            return true;
        };
        if (!monotone(P.group_vbegin, static_cast<std::uint64_t>(n_vsub)) ||
            !monotone(P.vsub_tbegin, static_cast<std::uint64_t>(n_terms)))
            throw qsa::Error(qsa::kInvalid, "program: offsets are not monotone or do not end at the counts");
        *out = new qedh_program_s{std::move(P), nullptr};
    });
}
int qedh_program_destroy(qedh_program p) {
    return guarded([&] { delete p; });
}

int qedh_sector_create(int n_sites, int n_up, int group_size, const int* perms, const uint64_t* flips,
                       const double* chars, int sublattice_mode, int relaxed, uint64_t rank_budget_bytes,
                       qedh_sector* out) {
    return guarded([&] {
        need(perms, "perms");
        need(chars, "chars");
        need(out, "out");
        if (n_sites < 1 || n_sites > 63 || group_size < 1) throw qsa::Error(qsa::kInvalid, "sector: bad sizes");
        qsa::GroupSpec g;
        g.n_sites = n_sites;
        g.n_up = n_up;
        g.G = group_size;
        g.perms.assign(perms, perms + static_cast<std::size_t>(group_size) * n_sites);
        if (flips) g.flips.assign(flips, flips + group_size);
        else g.flips.assign(group_size, 0ULL);
        g.chars.resize(group_size);
        for (int i = 0; i < group_size; ++i) g.chars[i] = qsa::cplx(chars[2 * i], chars[2 * i + 1]);
        g.slc = qsa::build_sublattice(g.perms, g.flips, group_size, n_sites, sublattice_mode, relaxed != 0);
        *out = new qedh_sector_s{std::make_shared<qsa::DeviceSector>(std::move(g), rank_budget_bytes)};
    });
}
int qedh_sector_create_irrep(int n_sites, int n_up, int group_size, const int* perms, const uint64_t* flips,
                             int irrep_dim, const double* D, int sublattice_mode, int relaxed,
                             uint64_t rank_budget_bytes, qedh_sector* out) {
    return guarded([&] {
        need(perms, "perms");
        need(D, "D");
        need(out, "out");
        if (n_sites < 1 || n_sites > 63 || group_size < 1 || irrep_dim < 1)
            throw qsa::Error(qsa::kInvalid, "sector: bad sizes");
        qsa::GroupSpec g;
        g.n_sites = n_sites;
        g.n_up = n_up;
        g.G = group_size;
        g.perms.assign(perms, perms + static_cast<std::size_t>(group_size) * n_sites);
        if (flips) g.flips.assign(flips, flips + group_size);
        else g.flips.assign(group_size, 0ULL);
        g.slc = qsa::build_sublattice(g.perms, g.flips, group_size, n_sites, sublattice_mode, relaxed != 0);
        const std::size_t n = static_cast<std::size_t>(group_size) * irrep_dim * irrep_dim;
        std::vector<qsa::cplx> Dm(n);
        for (std::size_t i = 0; i < n; ++i) Dm[i] = qsa::cplx(D[2 * i], D[2 * i + 1]);
        if (irrep_dim == 1) {   // a one-dimensional irrep: its characters
            g.chars = std::move(Dm);
            *out = new qedh_sector_s{std::make_shared<qsa::DeviceSector>(std::move(g), rank_budget_bytes)};
            return;
        }
        *out = new qedh_sector_s{qsa::DeviceSector::make_irrep(std::move(g), irrep_dim, Dm, rank_budget_bytes)};
    });
}
int qedh_sector_irrep_info(qedh_sector s, int* irrep_dim, int64_t* n_reps, int64_t* n_classes) {
    return guarded([&] {
        need(s, "sector");
        const auto& ir = s->sector->irrep();
        if (irrep_dim) *irrep_dim = ir.d;
        if (n_reps) *n_reps = s->sector->n_reps();
        if (n_classes) *n_classes = static_cast<int64_t>(ir.class_rank.size());
    });
}
int qedh_sector_irrep_export(qedh_sector s, uint16_t* rep_class, uint64_t* state_offset, uint8_t* class_rank,
                             double* class_C) {
    return guarded([&] {
        need(s, "sector");
        const auto& ir = s->sector->irrep();
        if (ir.d <= 1) throw qsa::Error(qsa::kInvalid, "irrep export: the sector's irrep is one-dimensional");
        if (rep_class) std::memcpy(rep_class, ir.rep_class.data(), ir.rep_class.size() * 2);
        if (state_offset) std::memcpy(state_offset, ir.state_offset.data(), ir.state_offset.size() * 8);
        if (class_rank) std::memcpy(class_rank, ir.class_rank.data(), ir.class_rank.size());
        if (class_C) std::memcpy(class_C, ir.class_C.data(), ir.class_C.size() * sizeof(qsa::cplx));
    });
}
int qedh_sector_import(int n_sites, int n_up, int group_size, const int* perms, const uint64_t* flips,
                       const double* chars, int sublattice_mode, int relaxed, int64_t n_reps, const uint64_t* reps,
                       const double* inv_norms, uint64_t rank_budget_bytes, qedh_sector* out) {
    return guarded([&] {
        need(perms, "perms");
        need(chars, "chars");
        need(out, "out");
        if (n_sites < 1 || n_sites > 63 || group_size < 1) throw qsa::Error(qsa::kInvalid, "sector: bad sizes");
        qsa::GroupSpec g;
        g.n_sites = n_sites;
        g.n_up = n_up;
        g.G = group_size;
        g.perms.assign(perms, perms + static_cast<std::size_t>(group_size) * n_sites);
        if (flips) g.flips.assign(flips, flips + group_size);
        else g.flips.assign(group_size, 0ULL);
        g.chars.resize(group_size);
        for (int i = 0; i < group_size; ++i) g.chars[i] = qsa::cplx(chars[2 * i], chars[2 * i + 1]);
        g.slc = qsa::build_sublattice(g.perms, g.flips, group_size, n_sites, sublattice_mode, relaxed != 0);
        *out = new qedh_sector_s{
            std::make_shared<qsa::DeviceSector>(std::move(g), reps, inv_norms, n_reps, rank_budget_bytes)};
    });
}
int qedh_sublattice_tables(int n_sites, int group_size, const int* perms, const uint64_t* flips, int sublattice_mode,
                           int relaxed, int* block, int* key_bit, uint16_t* lead, uint32_t* cand_off, uint16_t* cand,
                           int64_t cand_cap, int64_t* n_cand) {
    return guarded([&] {
        need(perms, "perms");
        need(block, "block");
        if (n_sites < 1 || n_sites > 63 || group_size < 1) throw qsa::Error(qsa::kInvalid, "sublattice: bad sizes");
        std::vector<int> P(perms, perms + static_cast<std::size_t>(group_size) * n_sites);
        std::vector<uint64_t> F = flips ? std::vector<uint64_t>(flips, flips + group_size)
                                        : std::vector<uint64_t>(static_cast<std::size_t>(group_size), 0ULL);
        const qsa::SublatticeTables t = qsa::build_sublattice(P, F, group_size, n_sites, sublattice_mode, relaxed != 0);
        *block = t.L;
        if (n_cand) *n_cand = static_cast<int64_t>(t.cand.size());
        if (!t.engaged()) return;
        if (key_bit) std::memcpy(key_bit, t.key_bit.data(), t.key_bit.size() * sizeof(int));
        if (lead) std::memcpy(lead, t.lead.data(), t.lead.size() * sizeof(uint16_t));
        if (cand_off) std::memcpy(cand_off, t.cand_off.data(), t.cand_off.size() * sizeof(uint32_t));
        if (cand) {
            if (cand_cap < static_cast<int64_t>(t.cand.size()))
                throw qsa::Error(qsa::kInvalid, "sublattice: cand_cap below the candidate count");
            std::memcpy(cand, t.cand.data(), t.cand.size() * sizeof(uint16_t));
        }
    });
}
int qedh_sector_info(qedh_sector s, int64_t* dim, int64_t* n_orbits, int* sublattice_block, double* build_seconds) {
    return guarded([&] {
        need(s, "sector");
        if (dim) *dim = s->sector->dim();
        if (n_orbits) *n_orbits = s->sector->stats().orbits;
        if (sublattice_block) *sublattice_block = s->sector->spec().slc.L;
        if (build_seconds) *build_seconds = s->sector->stats().seconds;
    });
}
int qedh_sector_export(qedh_sector s, uint64_t* reps, double* inv_norms) {
    return guarded([&] {
        need(s, "sector");
        if (reps) {
            const auto r = s->sector->reps();
            std::memcpy(reps, r.data(), r.size() * 8);
        }
        if (inv_norms) {
            const auto v = s->sector->inv_norms();
            std::memcpy(inv_norms, v.data(), v.size() * 8);
        }
    });
}
int qedh_sector_key_bits(qedh_sector s, int* key_bit, int* block) {
    return guarded([&] {
        need(s, "sector");
        const auto& slc = s->sector->spec().slc;
        if (block) *block = slc.L;
        if (key_bit && slc.engaged()) std::memcpy(key_bit, slc.key_bit.data(), slc.key_bit.size() * sizeof(int));
    });
}
int qedh_sector_destroy(qedh_sector s) {
    return guarded([&] { delete s; });
}

int qedh_matvec_create(qedh_sector rows, qedh_sector cols, qedh_program prog, double norm_bound, int lane,
                       uint64_t csr_budget_bytes, qedh_matvec* out) {
    return guarded([&] {
        need(rows, "rows");
        need(prog, "program");
        need(out, "out");
        auto c = (cols && cols != rows) ? cols->sector : rows->sector;
        const bool tune = (lane & QEDH_LANE_NO_TUNE) == 0;
        *out = new qedh_matvec_s{std::make_shared<qsa::SectorOperator>(rows->sector, c, device_program(prog), norm_bound,
                                                                      lane & ~QEDH_LANE_NO_TUNE, csr_budget_bytes,
                                                                      tune)};
    });
}
int qedh_matvec_info(qedh_matvec m, qedh_matvec_info_t* info) {
    return guarded([&] {
        need(m, "matvec");
        need(info, "info");
        const auto& o = *m->op;
        const auto& c = o.csr_info();
        info->rows = o.dim();
        info->cols = o.col_dim();
        info->is_real = o.is_real() ? 1 : 0;
        info->lane = o.lane();
        info->nnz = o.lane() == qsa::kLaneCsr ? c.nnz : 0;
        info->value_bytes = c.width;
        info->id_width_ref = static_cast<int>(c.id_width_ref);
        info->dict_size = c.dict;
        info->csr_bytes = c.bytes;
        info->build_seconds = c.build_seconds;
        info->lanes = m->op->sell() ? 1 : c.lanes;   // SELL-64: one lane per row
        info->norm_bound = o.norm_bound();
    });
}
int qedh_matvec_apply(qedh_matvec m, const void* in, void* out, int complex_vectors) {
    return guarded([&] {
        need(m, "matvec");
        m->op->apply(in, out, complex_vectors != 0);
    });
}
int qedh_matvec_apply_host(qedh_matvec m, const void* in, void* out, int complex_vectors) {
    return guarded([&] {
        need(m, "matvec");
        need(in, "in");
        need(out, "out");
        const auto& op = *m->op;
        const std::size_t es = (op.is_real() && !complex_vectors) ? 8 : 16;
        const std::size_t nin = static_cast<std::size_t>(op.col_dim()) * es, nout = static_cast<std::size_t>(op.dim()) * es;
        qsa::DeviceBuffer x(nin ? nin : 8), y(nout ? nout : 8);
        qsa::copy_to_device(x.get(), in, nin);
        op.apply(x.get(), y.get(), complex_vectors != 0);
        QSA_HIP(hipGetLastError());
        qsa::copy_to_host(out, y.get(), nout);
    });
}
int qedh_matvec_apply_multi(qedh_matvec m, const void* const* ins, void* const* outs, int k, int complex_vectors) {
    return guarded([&] {
        need(m, "matvec");
        m->op->apply_multi(ins, outs, k, complex_vectors != 0);
    });
}
int qedh_matvec_apply_block(qedh_matvec m, int nv, const void* X, void* Y, int complex_vectors) {
    return guarded([&] {
        need(m, "matvec");
        m->op->apply_block(nv, X, Y, complex_vectors != 0);
    });
}
int qedh_matvec_set_lanes(qedh_matvec m, int lanes) {
    return guarded([&] {
        need(m, "matvec");
        if (lanes != 4 && lanes != 8 && lanes != 16 && lanes != 32 && lanes != 64)
            throw qsa::Error(qsa::kInvalid, "lanes must be 4, 8, 16, 32 or 64");
        m->op->set_lanes(lanes);
    });
}
int qedh_matvec_export_csr(qedh_matvec m, uint64_t* row_ptr, uint32_t* col, double* val) {
    return guarded([&] {
        need(m, "matvec");
        std::vector<std::uint64_t> rp;
        std::vector<std::uint32_t> c;
        std::vector<double> v;
        m->op->csr_download(rp, c, v);
        if (row_ptr) std::memcpy(row_ptr, rp.data(), rp.size() * 8);
        if (col && !c.empty()) std::memcpy(col, c.data(), c.size() * 4);
        if (val && !v.empty()) std::memcpy(val, v.data(), v.size() * 8);
    });
}
int qedh_matvec_destroy(qedh_matvec m) {
    return guarded([&] { delete m; });
}

int qedh_lanczos(qedh_matvec m, const void* v0_host, int complex_vectors, int max_iter, int reorth,
                 double breakdown_tol, double* alpha, double* beta, int* iters, double* timings) {
    return guarded([&] {
        need(m, "matvec");
        need(v0_host, "v0");
        const auto& op = *m->op;
        const std::size_t n = static_cast<std::size_t>(op.dim());
        if (op.dim() != op.col_dim()) throw qsa::Error(qsa::kInvalid, "lanczos: the operator is not square");
        auto run = [&](auto tag) {
            using S = decltype(tag);
            qsa::HipBackend<S> be;
            auto v0 = be.make_zero_vector(n);
            be.copy_from_host(static_cast<const S*>(v0_host), v0.get(), n);
            qsa::LanczosOptions<S> o;
            o.max_iter = static_cast<std::size_t>(max_iter);
            o.reorth = static_cast<qsa::Reorth>(reorth);
            o.keep_basis = o.reorth == qsa::Reorth::FullCGS2;
            if (breakdown_tol > 0.0) o.breakdown_tol = breakdown_tol;
            const bool cv = qsa::HipBackend<S>::kComplex;
            auto r = qsa::lanczos_run(be, qsa::OperatorMatvec<S>{op, cv}, n, v0.get(), o);
            for (std::size_t j = 0; j < r.alpha.size(); ++j) alpha[j] = r.alpha[j];
            for (std::size_t j = 0; j < r.beta.size(); ++j) beta[j] = r.beta[j];
            *iters = static_cast<int>(r.iters);
            if (timings) {
                timings[0] = r.timings.apply_s;
                timings[1] = r.timings.reorth_s;
                timings[2] = r.timings.vector_s;
                timings[3] = r.timings.total_s;
            }
        };
        if (op.is_real() && !complex_vectors) run(double{});
        else run(qsa::cplx{});
    });
}

int qedh_block_lowest(qedh_matvec m, int k, int64_t max_iter, double* values, int* n_values, int* converged,
                      int64_t* applies) {
    return guarded([&] {
        need(m, "matvec");
        const auto& op = *m->op;
        const std::size_t n = static_cast<std::size_t>(op.dim());
        if (n == 0) {
            *n_values = 0;
            *converged = 1;
            return;
        }
        const std::size_t kk = std::max<std::size_t>(1, std::min<std::size_t>(static_cast<std::size_t>(k), n));
        auto run = [&](auto tag) {
            using S = decltype(tag);
            qsa::HipBackend<S> be;
            const bool cv = qsa::HipBackend<S>::kComplex;
            const qsa::OperatorMatvec<S> H{op, cv};
            qsa::BlockSolution sol;
            const std::uint64_t budget = max_iter > 0 ? static_cast<std::uint64_t>(max_iter) : 0;
            if (kk > 1) sol = qsa::krylov_schur_lane(be, H, n, kk, op.norm_bound(), budget);
            else sol = qsa::lowest_scan(be, H, n, budget > 0 ? budget : 400, op.norm_bound());
            *n_values = static_cast<int>(sol.values.size());
            for (std::size_t i = 0; i < sol.values.size(); ++i) values[i] = sol.values[i];
            *converged = sol.converged ? 1 : 0;
            if (applies) *applies = static_cast<int64_t>(sol.applies);
        };
        if (op.is_real()) run(double{});
        else run(qsa::cplx{});
    });
}

int qedh_block_estimate(qedh_matvec m, double* theta, double* residual, int64_t* applies) {
    return guarded([&] {
        need(m, "matvec");
        need(theta, "theta");
        need(residual, "residual");
        const auto& op = *m->op;
        if (!op.same()) throw qsa::Error(qsa::kInvalid, "estimate: the operator maps two sectors");
        const std::size_t n = static_cast<std::size_t>(op.dim());
        auto run = [&](auto tag) {
            using S = decltype(tag);
            qsa::HipBackend<S> be;
            const bool cv = qsa::HipBackend<S>::kComplex;
            const qsa::OperatorMatvec<S> H{op, cv};
            const qsa::BlockEstimate e = qsa::estimate_lowest(be, H, n);
            *theta = e.theta;
            *residual = e.residual;
            if (applies) *applies = static_cast<int64_t>(e.applies);
        };
        if (op.is_real()) run(double{});
        else run(qsa::cplx{});
    });
}

int qedh_block_eigenpairs(qedh_matvec m, int k, int64_t max_iter, double* values, void* vectors, int* n_values,
                          int* converged) {
    return guarded([&] {
        need(m, "matvec");
        need(values, "values");
        need(vectors, "vectors");
        const auto& op = *m->op;
        if (!op.same()) throw qsa::Error(qsa::kInvalid, "eigenpairs: the operator maps two sectors");
        const std::size_t n = static_cast<std::size_t>(op.dim());
        if (k < 1) throw qsa::Error(qsa::kInvalid, "eigenpairs: k must be >= 1");
        const std::size_t kk = std::min<std::size_t>(static_cast<std::size_t>(k), n);
        qsa::HipBackend<qsa::cplx> be;
        const qsa::OperatorMatvec<qsa::cplx> H{op, true};
        const qsa::BlockSolution sol =
            qsa::krylov_schur_lane(be, H, n, kk, op.norm_bound(), max_iter > 0 ? static_cast<std::uint64_t>(max_iter) : 0,
                                   static_cast<qsa::cplx*>(vectors));
        be.synchronize();
        *n_values = static_cast<int>(sol.values.size());
        for (std::size_t i = 0; i < sol.values.size(); ++i) values[i] = sol.values[i];
        if (converged) *converged = sol.converged ? 1 : 0;
    });
}

int qedh_ftlm(qedh_matvec m, int samples, int krylov, uint64_t seed, int n_beta, const double* betas,
              double breakdown_tol, double* lnZ, double* E, double* V, double* e_min) {
    return guarded([&] {
        need(m, "matvec");
        const auto& op = *m->op;
        const std::size_t n = static_cast<std::size_t>(op.dim());
        if (samples < 1 || krylov < 2 || n_beta < 1) throw qsa::Error(qsa::kInvalid, "ftlm: bad sizes");
        qsa::HipBackend<qsa::cplx> be;
        qsa::FtlmOptions o;
        o.samples = static_cast<std::size_t>(samples);
        o.krylov = static_cast<std::size_t>(krylov);
        o.seed = seed;
        o.betas.assign(betas, betas + n_beta);
        o.breakdown_tol = breakdown_tol;
        const auto c = qsa::ftlm_block(be, qsa::OperatorMatvec<qsa::cplx>{op, true}, n, o);
        for (int t = 0; t < n_beta; ++t) {
            lnZ[t] = c.lnZ[t];
            E[t] = c.E[t];
            V[t] = c.V[t];
        }
        if (e_min) *e_min = c.e_min;
    });
}

int qedh_cf(qedh_matvec m, const double* phi_host, int krylov, int n_omega, const double* omega, double eta,
            int has_shift, double energy_shift, double breakdown_tol, double* S, double* out_shift, double* out_norm_sq,
            double* out_change) {
    return guarded([&] {
        need(m, "matvec");
        need(phi_host, "phi");
        const auto& op = *m->op;
        const std::size_t n = static_cast<std::size_t>(op.dim());
        qsa::HipBackend<qsa::cplx> be;
        auto phi = be.make_zero_vector(n);
        be.copy_from_host(reinterpret_cast<const qsa::cplx*>(phi_host), phi.get(), n);
        std::vector<double> w(omega, omega + n_omega);
        const auto r = qsa::cf_from_vector(
            be, qsa::OperatorMatvec<qsa::cplx>{op, true}, n, phi.get(), w, eta,
            static_cast<std::size_t>(krylov), has_shift != 0, energy_shift, breakdown_tol > 0.0 ? breakdown_tol : 1e-12);
        for (int i = 0; i < n_omega; ++i) S[i] = r.S[i];
        if (out_shift) *out_shift = r.energy_shift;
        if (out_norm_sq) *out_norm_sq = r.norm_sq;
        if (out_change) *out_change = r.convergence_change;
    });
}

}  // extern "C"

extern "C" int qedh_matrix_elements(qedh_sector bra_sector, qedh_sector ket_sector, int n_obs,
                                    const qedh_program* programs, int n_pairs, const void* const* bras,
                                    const void* const* kets, double* out) {
    return guarded([&] {
        need(bra_sector, "bra sector");
        need(ket_sector, "ket sector");
        need(programs, "programs");
        need(bras, "bras");
        need(kets, "kets");
        need(out, "out");
        if (n_obs < 1 || n_pairs < 1 || n_pairs > 4096)
            throw qsa::Error(qsa::kInvalid, "matrix elements: 1 or more observables and 1 to 4096 pairs");
        const auto& B = *bra_sector->sector;
        const auto& K = *ket_sector->sector;
        if (B.spec().perms != K.spec().perms || B.spec().flips != K.spec().flips)
            throw qsa::Error(qsa::kInvalid, "matrix elements: sectors of different groups");
        if (B.irrep_dim() > 1 || K.irrep_dim() > 1)
            throw qsa::Error(qsa::kDeviceUnsupported, "matrix elements: sectors of irreps of dimension > 1");
        if (B.spec().slc.engaged() != K.spec().slc.engaged() || B.spec().slc.key_bit != K.spec().slc.key_bit)
            throw qsa::Error(qsa::kInvalid, "matrix elements: the sectors use different representative orders");
        std::vector<const qsa::Program*> ps;
        for (int o = 0; o < n_obs; ++o) {
            need(programs[o], "program");
            ps.push_back(&programs[o]->prog);
        }
        for (int p = 0; p < n_pairs; ++p) {
            need(bras[p], "bra vector");
            need(kets[p], "ket vector");
        }
        const qsa::MultiProgram M = qsa::merge_programs(ps);
        const qsa::DeviceProgram dp(M.prog);
        const qsa::DeviceBuffer tobs = qsa::upload(M.term_obs.data(), M.term_obs.size());
        // per flip group its observables, ascending (the list every lane of a wave walks together)
        std::vector<std::uint32_t> gbegin(1, 0), gobs;
        for (std::size_t g = 0; g + 1 < M.prog.group_vbegin.size(); ++g) {
            const std::uint32_t k0 = M.prog.vsub_tbegin[M.prog.group_vbegin[g]];
            const std::uint32_t k1 = M.prog.vsub_tbegin[M.prog.group_vbegin[g + 1]];
            std::vector<std::uint32_t> os(M.term_obs.begin() + k0, M.term_obs.begin() + k1);
            std::sort(os.begin(), os.end());
            os.erase(std::unique(os.begin(), os.end()), os.end());
            gobs.insert(gobs.end(), os.begin(), os.end());
            gbegin.push_back(static_cast<std::uint32_t>(gobs.size()));
        }
        if (gobs.empty()) gobs.push_back(0);
        const qsa::DeviceBuffer gb = qsa::upload(gbegin.data(), gbegin.size());
        const qsa::DeviceBuffer go = qsa::upload(gobs.data(), gobs.size());
        const qsa::DeviceBuffer bd = qsa::upload(bras, static_cast<std::size_t>(n_pairs));
        const qsa::DeviceBuffer kd = qsa::upload(kets, static_cast<std::size_t>(n_pairs));
        const int pc_max = std::min(n_pairs, qsa::dev::kObsCells);
        const int blocks = qsa::dev::obs_sweep_blocks(B.n_reps());
        const int cells_max = std::min(n_pairs * n_obs, qsa::dev::kObsCells);
        qsa::DeviceBuffer partial(static_cast<std::size_t>(blocks) * 2 * cells_max * sizeof(double));
        qsa::DeviceBuffer dout(static_cast<std::size_t>(n_pairs) * n_obs * 2 * sizeof(double));
        const hipStream_t st = qsa::engine_stream();
        for (int p0 = 0; p0 < n_pairs; p0 += pc_max) {
            const int pc = std::min(pc_max, n_pairs - p0);
            const int nob = std::min(n_obs, qsa::dev::kObsCells / pc);
            for (int o0 = 0; o0 < n_obs; o0 += nob)
                qsa::dev::obs_sweep(B.view(), K.view(), bra_sector->sector == ket_sector->sector, dp.view(),
                                    tobs.get<std::uint32_t>(), gb.get<std::uint32_t>(), go.get<std::uint32_t>(), o0,
                                    std::min(nob, n_obs - o0), n_obs, pc,
                                    bd.get<const qsa::dev::Zd* const>() + p0, kd.get<const qsa::dev::Zd* const>() + p0,
                                    partial.get<double>(), dout.get<double>() + static_cast<std::size_t>(p0) * n_obs * 2,
                                    st);
        }
        QSA_HIP(hipGetLastError());
        qsa::copy_to_host(out, dout.get(), dout.bytes());
    });
}
extern "C" int qedh_sector_index_and_projection(qedh_sector s, const uint64_t* states, int64_t n, int64_t* index,
                                                double* proj) {
    return guarded([&] {
        need(s, "sector");
        if (s->sector->irrep_dim() > 1)
            throw qsa::Error(qsa::kDeviceUnsupported, "index_and_projection: the sector's irrep has dimension > 1");
        if (n <= 0) return;
        auto st = qsa::upload(states, static_cast<std::size_t>(n));
        qsa::DeviceBuffer di(static_cast<std::size_t>(n) * 8), dp(static_cast<std::size_t>(n) * 16);
        qsa::dev::index_proj_batch(s->sector->view(), st.get<std::uint64_t>(), n, di.get<qsa::dev::i64>(),
                                   dp.get<double>(), qsa::engine_stream());
        QSA_HIP(hipGetLastError());
        qsa::copy_to_host(index, di.get(), static_cast<std::size_t>(n) * 8);
        qsa::copy_to_host(proj, dp.get(), static_cast<std::size_t>(n) * 16);
    });
}

extern "C" int qedh_sector_expand(qedh_sector s, const void* u, int nvec, int full, void* out, double* norms) {
    return guarded([&] {
        need(s, "sector");
        need(u, "u");
        need(out, "out");
        if (nvec < 1) throw qsa::Error(qsa::kInvalid, "expand: nvec must be >= 1");
        if (s->sector->irrep_dim() > 1)
            throw qsa::Error(qsa::kDeviceUnsupported, "expand: sectors of irreps of dimension > 1 are not expanded");
        const qsa::dev::SectorView S = s->sector->view();
        const int N = S.n_sites;
        if (full && N > 34)
            throw qsa::Error(qsa::kResourceLimit, "expand: the full 2^N space is limited to N <= 34 (use the Sz basis)");
        if (!full && S.n_up < 0) throw qsa::Error(qsa::kInvalid, "expand: a sector without fixed n_up expands only into the full space");
        std::uint64_t out_dim = 1;
        if (full) {
            out_dim = std::uint64_t{1} << N;
        } else {
            for (int i = 1; i <= S.n_up; ++i)
                out_dim = out_dim * static_cast<std::uint64_t>(N - S.n_up + i) / static_cast<std::uint64_t>(i);
        }
        const hipStream_t st = qsa::engine_stream();
        qsa::DeviceBuffer bad(sizeof(unsigned)), scratch(qsa::dev::blas_scratch_bytes(1)),
            nrm(static_cast<std::size_t>(nvec) * sizeof(double));
        QSA_HIP(hipMemsetAsync(bad.get(), 0, sizeof(unsigned), st));
        const auto* uin = static_cast<const qsa::dev::Zd*>(u);
        auto* psi = static_cast<qsa::dev::Zd*>(out);
        for (int v = 0; v < nvec; ++v) {
            qsa::dev::Zd* p = psi + static_cast<std::size_t>(v) * out_dim;
            qsa::dev::fill_zero(true, p, out_dim, st);
            qsa::dev::sector_expand(S, uin + static_cast<std::size_t>(v) * static_cast<std::size_t>(S.dim), full != 0,
                                    p, bad.get<unsigned>(), st);
            qsa::dev::nrm2(true, p, out_dim, scratch.get(), nrm.get<double>() + v, st);
        }
        QSA_HIP(hipGetLastError());
        unsigned bad_h = 0;
        std::vector<double> nh(static_cast<std::size_t>(nvec));
        qsa::copy_to_host(&bad_h, bad.get(), sizeof(unsigned));
        qsa::copy_to_host(nh.data(), nrm.get(), nh.size() * sizeof(double));
        if (bad_h) throw qsa::Error(qsa::kInvalid, "expand: the sector is not inside the Sz sector n_up");
        for (int v = 0; v < nvec; ++v) {
            if (!(nh[static_cast<std::size_t>(v)] > 0.0))
                throw qsa::Error(qsa::kInvalid, "expand: the expansion annihilated the vector");
            qsa::dev::scale(true, qsa::dev::Zd{1.0 / nh[static_cast<std::size_t>(v)], 0.0},
                            psi + static_cast<std::size_t>(v) * out_dim, out_dim, st);
            if (norms) norms[v] = nh[static_cast<std::size_t>(v)];
        }
        QSA_HIP(hipGetLastError());
        QSA_HIP(hipStreamSynchronize(st));
    });
}

namespace qsa {
void capi_set_error(const std::string& m) { g_error = m; }
int capi_status(const std::exception& e) {
    g_error = e.what();
    return kRuntime;
}
}  // namespace qsa
