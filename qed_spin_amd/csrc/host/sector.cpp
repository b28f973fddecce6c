#include "sector.h"

#include <array>
#include <cmath>
#include <map>
#include <mutex>

#include <chrono>

#include "../device/launch.h"
#include "algebra.h"

namespace qsa {

namespace {
std::uint64_t binom_host(int n, int k) {
    if (k < 0 || n < 0 || k > n) return 0;
    std::uint64_t r = 1;
    for (int i = 1; i <= k; ++i) r = r * static_cast<std::uint64_t>(n - k + i) / static_cast<std::uint64_t>(i);
    return r;
}
}  // namespace

void device_alloc(void** p, std::size_t bytes) {
    *p = nullptr;
    int dev = 0;
    QSA_HIP(hipGetDevice(&dev));
    static std::once_flag once[64];
    if (dev >= 0 && dev < 64)
        std::call_once(once[dev], [dev] {
            hipMemPool_t pool = nullptr;
            if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess && pool) {
                std::uint64_t keep = ~std::uint64_t{0};
                (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
            }
            (void)hipGetLastError();
        });
    const hipError_t e = hipMallocAsync(p, bytes, engine_stream());
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        throw Error(e == hipErrorOutOfMemory ? kResourceLimit : kRuntime,
                    "device allocation of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
    }
}

void device_free(void* p) noexcept {
    if (p) (void)hipFreeAsync(p, engine_stream());
}

void require_device() {
    int n = 0;
    const hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        (void)hipGetLastError();
        throw Error(kDeviceUnavailable, "no HIP device is visible");
    }
}

void DeviceSector::setup_tables() {
    require_device();
    const int N = spec_.n_sites, G = spec_.G;
    if (N < 1 || N > 63) throw Error(kInvalid, "sector: n_sites must be in [1, 63]");
    if (G < 1) throw Error(kInvalid, "sector: the group needs at least the identity");
    if (static_cast<int>(spec_.perms.size()) != G * N || static_cast<int>(spec_.chars.size()) != G)
        throw Error(kInvalid, "sector: permutations / characters do not match the group size");
    if (spec_.flips.empty()) spec_.flips.assign(G, 0ULL);
    if (spec_.n_up > N) throw Error(kInvalid, "sector: n_up > n_sites");
    if (spec_.n_up < 0 && N > 40) throw Error(kDeviceUnsupported, "sector: a full-space sector above 40 sites");
    for (int g = 0; g < G; ++g) {
        std::uint64_t seen = 0;
        for (int i = 0; i < N; ++i) {
            const int p = spec_.perms[static_cast<std::size_t>(g) * N + i];
            if (p < 0 || p >= N) throw Error(kInvalid, "sector: permutation entry out of range");
            seen |= 1ULL << p;
        }
        if (seen != ((1ULL << N) - 1ULL)) throw Error(kInvalid, "sector: an element is not a permutation");
        if (spec_.n_up >= 0 && spec_.flips[g] != 0 && 2 * spec_.n_up != N)
            throw Error(kInvalid, "sector: a flip element needs n_up = N/2");
        if ((spec_.flips[g] & ~((1ULL << N) - 1ULL)) != 0) throw Error(kInvalid, "sector: flip mask beyond n_sites");
        if (spec_.n_up >= 0 && spec_.flips[g] != 0 && spec_.flips[g] != ((1ULL << N) - 1ULL))
            throw Error(kInvalid, "sector: a fixed-Sz sector takes only the global spin flip");
    }

    const auto lut = spec_.byte_lut();
    lut_ = upload(lut.data(), lut.size());
    flips_ = upload(spec_.flips.data(), spec_.flips.size());
    std::vector<double> cr(G), ci(G);
    for (int g = 0; g < G; ++g) {
        cr[g] = spec_.chars[g].real();
        ci[g] = spec_.chars[g].imag();
    }
    chi_re_ = upload(cr.data(), cr.size());
    chi_im_ = upload(ci.data(), ci.size());
    std::vector<std::uint64_t> pascal(65 * 65, 0);
    for (int n = 0; n <= 64; ++n)
        for (int k = 0; k <= n; ++k) pascal[n * 65 + k] = binom_host(n, k);
    binom_ = upload(pascal.data(), pascal.size());
    {
        const std::vector<std::uint64_t> nib = spec_.nibble_tables();
        const std::size_t body = static_cast<std::size_t>(G) * spec_.nq() * 16;
        nib_ = upload(nib.data(), nib.size());
        nib_sizes_ = {body, static_cast<std::size_t>(G), static_cast<std::size_t>(spec_.nq()) * 16};
    }
    if (spec_.slc.engaged()) {
        to_key_ = upload(spec_.slc.to_key.data(), spec_.slc.to_key.size());
        // one record per (block, pattern): lead | count << 16 | (the element, or the list's offset) << 32
        const auto& t = spec_.slc;
        std::vector<std::uint64_t> rec(t.lead.size());
        for (std::size_t e = 0; e < rec.size(); ++e) {
            const std::uint32_t off = t.cand_off[e], cnt = t.cand_off[e + 1] - off;
            // (a block no element maps to the leading one has no candidates: its lead is the 0xFFFF
            // sentinel and its empty list adds nothing)
            if (cnt > 0xFFFFu) throw Error(kDeviceUnsupported, "sector: more than 65535 candidate elements");
            const std::uint64_t pay = cnt == 1 ? t.cand[off] : off;
            rec[e] = static_cast<std::uint64_t>(t.lead[e]) | (static_cast<std::uint64_t>(cnt) << 16) | (pay << 32);
        }
        cand_rec_ = upload(rec.data(), rec.size());
        cand_ = upload(spec_.slc.cand.data(), spec_.slc.cand.size());
    }
}

void DeviceSector::build_rank_dir(std::uint64_t total, std::uint64_t rank_budget_bytes, const std::uint64_t* kept,
                                  const std::uint64_t* offsets) {
    const hipStream_t st = engine_stream();
    const std::uint64_t n_words = (total + 63) / 64;
    const std::uint64_t dir_bytes = n_words * sizeof(dev::RankRec);
    if (dim_ > 0 && dir_bytes <= rank_budget_bytes) {
        rank_dir_.reset(dir_bytes);
        if (kept && offsets) {
            dev::rank_pack(kept, offsets, n_words, rank_dir_.get<dev::RankRec>(), st);
        } else {
            // imported representatives: their bitmap, its scan, then the records
            DeviceBuffer bits(n_words * 8), counts(n_words * 8), offs(n_words * 8);
            QSA_HIP(hipMemsetAsync(bits.get(), 0, n_words * 8, st));
            dev::SectorView v = view();
            v.rank_dir = nullptr;
            dev::rank_mark(v, reps_.get<std::uint64_t>(), dim_, bits.get<std::uint64_t>(), st);
            QSA_HIP(hipGetLastError());
            const std::size_t tb = dev::scan_words_temp_bytes(n_words);
            DeviceBuffer temp(tb ? tb : 8);
            dev::scan_words(bits.get<std::uint64_t>(), n_words, counts.get<std::uint64_t>(),
                            offs.get<std::uint64_t>(), temp.get(), tb, st);
            dev::rank_pack(bits.get<std::uint64_t>(), offs.get<std::uint64_t>(), n_words,
                           rank_dir_.get<dev::RankRec>(), st);
            QSA_HIP(hipStreamSynchronize(st));
        }
        QSA_HIP(hipGetLastError());
        stats_.rank_table = true;
    }
    QSA_HIP(hipStreamSynchronize(st));
    stats_.device_bytes = lut_.bytes() + nib_.bytes() + reps_.bytes() + inv_norms_.bytes() + rank_dir_.bytes() + cand_rec_.bytes() +
                          cand_.bytes();
}

DeviceSector::DeviceSector(GroupSpec spec, std::uint64_t rank_budget_bytes) : spec_(std::move(spec)) {
    const auto t0 = std::chrono::steady_clock::now();
    setup_tables();
    const int N = spec_.n_sites;

    const std::uint64_t total = spec_.n_up >= 0 ? binom_host(N, spec_.n_up) : (1ULL << N);
    stats_.subspace_dim = total;
    const std::uint64_t n_words = (total + 63) / 64;
    if (n_words > 0x7FFFFFFFULL) throw Error(kDeviceUnsupported, "sector: subspace beyond 2^37 states");
    const hipStream_t st = engine_stream();
    dev::SectorView v = view();
    DeviceBuffer kept(n_words * 8), counts(n_words * 8), offsets(n_words * 8), orbits(8);
    QSA_HIP(hipMemsetAsync(orbits.get(), 0, 8, st));
    dev::orbit_flags(v, total, kept.get<std::uint64_t>(), orbits.get<unsigned long long>(), st);
    QSA_HIP(hipGetLastError());
    const std::size_t tb = dev::scan_words_temp_bytes(n_words);
    DeviceBuffer temp(tb ? tb : 8);
    dev::scan_words(kept.get<std::uint64_t>(), n_words, counts.get<std::uint64_t>(), offsets.get<std::uint64_t>(),
                    temp.get(), tb, st);
    QSA_HIP(hipGetLastError());
    std::uint64_t last[2] = {0, 0};
    unsigned long long n_orbits = 0;
    QSA_HIP(hipMemcpyAsync(&last[0], offsets.get<std::uint64_t>() + (n_words - 1), 8, hipMemcpyDeviceToHost, st));
    QSA_HIP(hipMemcpyAsync(&last[1], counts.get<std::uint64_t>() + (n_words - 1), 8, hipMemcpyDeviceToHost, st));
    QSA_HIP(hipMemcpyAsync(&n_orbits, orbits.get(), 8, hipMemcpyDeviceToHost, st));
    QSA_HIP(hipStreamSynchronize(st));
    dim_ = static_cast<std::int64_t>(last[0] + last[1]);
    stats_.orbits = static_cast<std::int64_t>(n_orbits);
    if (dim_ > 0x7FFFFFFFLL) throw Error(kDeviceUnsupported, "sector: more than 2^31 representatives");
    reps_.reset(static_cast<std::size_t>(dim_) * 8);
    inv_norms_.reset(static_cast<std::size_t>(dim_) * 8);
    if (dim_ > 0) {
        dev::orbit_emit(v, total, kept.get<std::uint64_t>(), offsets.get<std::uint64_t>(), reps_.get<std::uint64_t>(),
                        inv_norms_.get<double>(), st);
        QSA_HIP(hipGetLastError());
    }
    build_rank_dir(total, rank_budget_bytes, kept.get<std::uint64_t>(), offsets.get<std::uint64_t>());
    stats_.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

DeviceSector::DeviceSector(GroupSpec spec, const std::uint64_t* reps, const double* inv_norms, std::int64_t n,
                           std::uint64_t rank_budget_bytes)
    : spec_(std::move(spec)) {
    const auto t0 = std::chrono::steady_clock::now();
    setup_tables();
    const int N = spec_.n_sites;
    if (n < 0 || (n > 0 && (!reps || !inv_norms))) throw Error(kInvalid, "sector import: no representatives");
    if (n > 0x7FFFFFFFLL) throw Error(kDeviceUnsupported, "sector import: more than 2^31 representatives");
    const std::uint64_t total = spec_.n_up >= 0 ? binom_host(N, spec_.n_up) : (1ULL << N);
    const std::uint64_t mask = N == 64 ? ~0ULL : ((1ULL << N) - 1ULL);
    for (std::int64_t i = 0; i < n; ++i) {
        if ((reps[i] & ~mask) != 0 || (spec_.n_up >= 0 && __builtin_popcountll(reps[i]) != spec_.n_up))
            throw Error(kInvalid, "sector import: representative " + std::to_string(i) + " is not a state of the subspace");
        if (!(inv_norms[i] > 0.0)) throw Error(kInvalid, "sector import: 1/norm " + std::to_string(i) + " is not positive");
    }
    // ascending states (the lookup's order; with the sublattice code engaged each representative is
    // its orbit's least image in key order, but the table stays in state order)
    for (std::int64_t i = 1; i < n; ++i)
        if (!(reps[i - 1] < reps[i]))
            throw Error(kInvalid, "sector import: representative " + std::to_string(i) +
                                      " is out of order (the lookup needs them strictly ascending)");
    stats_.subspace_dim = total;
    stats_.orbits = -1;   // not scanned
    dim_ = n;
    reps_ = upload(reps, static_cast<std::size_t>(n));
    inv_norms_ = upload(inv_norms, static_cast<std::size_t>(n));
    build_rank_dir(total, rank_budget_bytes, nullptr, nullptr);
    // every representative must be its own orbit's representative and find itself: the order
    // (plain or key order) and the canonical form are the ones this group's lookup uses
    if (n > 0) {
        const hipStream_t st = engine_stream();
        DeviceBuffer idx(static_cast<std::size_t>(n) * 8), proj(static_cast<std::size_t>(n) * 16);
        dev::index_proj_batch(view(), reps_.get<std::uint64_t>(), n, idx.get<dev::i64>(), proj.get<double>(), st);
        QSA_HIP(hipGetLastError());
        std::vector<std::int64_t> ih(static_cast<std::size_t>(n));
        copy_to_host(ih.data(), idx.get(), ih.size() * 8);
        for (std::int64_t i = 0; i < n; ++i)
            if (ih[static_cast<std::size_t>(i)] != i)
                throw Error(kInvalid, "sector import: representative " + std::to_string(i) +
                                          " is not its orbit's representative in this group's order (it maps to " +
                                          std::to_string(ih[static_cast<std::size_t>(i)]) + ")");
    }
    stats_.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

dev::SectorView DeviceSector::view() const {
    dev::SectorView v;
    v.n_sites = spec_.n_sites;
    v.n_up = spec_.n_up;
    v.G = spec_.G;
    v.bpw = spec_.bpw();
    v.lut = lut_.get<std::uint64_t>();
    v.flips = flips_.get<std::uint64_t>();
    v.chi_re = chi_re_.get<double>();
    v.chi_im = chi_im_.get<double>();
    v.nq = spec_.nq();
    v.nib = nib_.get<std::uint64_t>();
    v.nib_flip = v.nib + nib_sizes_[0];
    if (spec_.slc.engaged()) {
        v.nib_key = v.nib_flip + nib_sizes_[1];
        v.nib_unkey = v.nib_key + nib_sizes_[2];
    }
    if (spec_.slc.engaged()) {
        v.slc_L = spec_.slc.L;
        v.slc_m = spec_.slc.m;
        v.to_key = to_key_.get<std::uint64_t>();
        v.cand_rec = cand_rec_.get<std::uint64_t>();
        v.cand = cand_.get<std::uint16_t>();
    }
    v.reps = reps_.get<std::uint64_t>();
    v.inv_norms = inv_norms_.get<double>();
    v.dim = dim_;
    v.rank_dir = rank_dir_ ? rank_dir_.get<dev::RankRec>() : nullptr;
    v.binom = binom_.get<std::uint64_t>();
    return v;
}

std::vector<std::uint64_t> DeviceSector::reps() const {
    std::vector<std::uint64_t> h(static_cast<std::size_t>(dim_));
    if (dim_) copy_to_host(h.data(), reps_.get(), h.size() * 8);
    return h;
}

std::vector<double> DeviceSector::inv_norms() const {
    std::vector<double> h(static_cast<std::size_t>(dim_));
    if (dim_) copy_to_host(h.data(), inv_norms_.get(), h.size() * 8);
    return h;
}

namespace {

/// An orthonormal basis of the range of the Hermitian projector P (d x d, row-major): Gram-Schmidt
/// (twice) over its columns in order, a column kept when a part of norm above 1e-6 remains. The
/// basis vectors are the columns of U (d x d, row-major; columns from the rank on zero).
int projector_range(const std::vector<cplx>& P, int d, std::vector<cplx>& U) {
    const std::size_t dd = static_cast<std::size_t>(d) * static_cast<std::size_t>(d);
    U.assign(dd, cplx(0.0, 0.0));
    int rank = 0;
    std::vector<cplx> v(static_cast<std::size_t>(d));
    for (int j = 0; j < d; ++j) {
        for (int i = 0; i < d; ++i) v[i] = P[static_cast<std::size_t>(i) * d + j];
        for (int pass = 0; pass < 2; ++pass)
            for (int k = 0; k < rank; ++k) {
                cplx c(0.0, 0.0);
                for (int i = 0; i < d; ++i) c += std::conj(U[static_cast<std::size_t>(i) * d + k]) * v[i];
                for (int i = 0; i < d; ++i) v[i] -= c * U[static_cast<std::size_t>(i) * d + k];
            }
        double n2 = 0.0;
        for (int i = 0; i < d; ++i) n2 += std::norm(v[i]);
        if (!(n2 > 1e-12)) continue;
        const double inv = 1.0 / std::sqrt(n2);
        for (int i = 0; i < d; ++i) U[static_cast<std::size_t>(i) * d + rank] = v[i] * inv;
        ++rank;
    }
    return rank;
}

}  // namespace

std::shared_ptr<DeviceSector> DeviceSector::make_irrep(GroupSpec spec, int d, const std::vector<cplx>& D,
                                                       std::uint64_t rank_budget_bytes) {
    const int G = spec.G;
    if (d < 2) throw Error(kInvalid, "irrep sector: an irrep of dimension at least 2 (use the characters for d = 1)");
    if (d > dev::kMaxIrrepDim)
        throw Error(kDeviceUnsupported, "irrep sector: irreps up to dimension " + std::to_string(dev::kMaxIrrepDim) +
                                            " (the reference's limit)");
    if (G < 1 || G > 256) throw Error(kDeviceUnsupported, "irrep sector: groups of 1 to 256 elements");
    const std::size_t dd = static_cast<std::size_t>(d) * static_cast<std::size_t>(d);
    if (D.size() != static_cast<std::size_t>(G) * dd) throw Error(kInvalid, "irrep sector: D needs G x d x d entries");
    for (const cplx& z : D)
        if (!std::isfinite(z.real()) || !std::isfinite(z.imag())) throw Error(kInvalid, "irrep sector: D is not finite");
    spec.chars.assign(static_cast<std::size_t>(G), cplx(0.0, 0.0));   // the characters label the sector
    for (int g = 0; g < G; ++g)
        for (int i = 0; i < d; ++i) spec.chars[g] += D[static_cast<std::size_t>(g) * dd + static_cast<std::size_t>(i) * (d + 1)];
    // every orbit's representative: the trivial characters keep them all
    GroupSpec trivial = spec;
    trivial.chars.assign(static_cast<std::size_t>(G), cplx(1.0, 0.0));
    std::vector<std::uint64_t> reps;
    std::vector<std::uint64_t> masks;
    const int words = (G + 63) / 64;
    std::int64_t n = 0, orbits = 0;
    {
        const DeviceSector all(std::move(trivial), 0);
        n = all.dim_;
        orbits = all.stats_.orbits;
        reps = all.reps();
        masks.assign(static_cast<std::size_t>(n) * words, 0);
        if (n > 0) {
            DeviceBuffer m(static_cast<std::size_t>(n) * words * 8);
            dev::stab_masks(all.view(), words, m.get<std::uint64_t>(), engine_stream());
            QSA_HIP(hipGetLastError());
            copy_to_host(masks.data(), m.get(), masks.size() * 8);
        }
    }
    // stabiliser classes: P = (1 / |S|) sum_{s in S} conj(D(s)) is a projector; C = an orthonormal
    // basis of its range over sqrt(|S|), so C^dag (|S| P) C = I
    IrrepData ir;
    ir.d = d;
    ir.D = D;
    std::map<std::array<std::uint64_t, 4>, int> class_of;
    std::vector<int> cls(static_cast<std::size_t>(n));
    std::array<std::uint64_t, 4> last{};
    int last_class = -1;
    for (std::int64_t r = 0; r < n; ++r) {
        std::array<std::uint64_t, 4> key{};
        for (int w = 0; w < words; ++w) key[w] = masks[static_cast<std::size_t>(r) * words + w];
        if (last_class >= 0 && key == last) {
            cls[r] = last_class;
            continue;
        }
        auto it = class_of.find(key);
        int c;
        if (it != class_of.end()) {
            c = it->second;
        } else {
            c = static_cast<int>(class_of.size());
            if (c > 65535) throw Error(kDeviceUnsupported, "irrep sector: more than 65536 stabiliser classes");
            class_of.emplace(key, c);
            std::vector<cplx> P(dd, cplx(0.0, 0.0));
            int size = 0;
            for (int g = 0; g < G; ++g)
                if ((key[g / 64] >> (g % 64)) & 1ULL) {
                    ++size;
                    for (std::size_t e = 0; e < dd; ++e) P[e] += std::conj(D[static_cast<std::size_t>(g) * dd + e]);
                }
            if (size == 0) throw Error(kRuntime, "irrep sector: a representative with an empty stabiliser");
            std::vector<cplx> H(dd);
            for (int i = 0; i < d; ++i)
                for (int j = 0; j < d; ++j)
                    H[static_cast<std::size_t>(i) * d + j] =
                        0.5 * (P[static_cast<std::size_t>(i) * d + j] + std::conj(P[static_cast<std::size_t>(j) * d + i])) /
                        static_cast<double>(size);
            std::vector<cplx> U;
            const int rank = projector_range(H, d, U);
            double tr = 0.0;
            for (int i = 0; i < d; ++i) tr += H[static_cast<std::size_t>(i) * (d + 1)].real();
            if (std::abs(tr - rank) > 1e-6)
                throw Error(kInvalid, "irrep sector: D is not a unitary representation of the group (a stabiliser's "
                                      "projector has trace " + std::to_string(tr) + " and rank " + std::to_string(rank) + ")");
            const double sc = 1.0 / std::sqrt(static_cast<double>(size));
            for (cplx& z : U) z *= sc;
            ir.class_C.insert(ir.class_C.end(), U.begin(), U.end());
            ir.class_rank.push_back(static_cast<std::uint8_t>(rank));
        }
        cls[r] = c;
        last = key;
        last_class = c;
    }
    // the representatives that hold states of the irrep
    std::vector<std::uint64_t> keep;
    ir.state_offset.assign(1, 0);
    for (std::int64_t r = 0; r < n; ++r) {
        const int rank = ir.class_rank[static_cast<std::size_t>(cls[r])];
        if (rank == 0) continue;
        keep.push_back(reps[static_cast<std::size_t>(r)]);
        ir.rep_class.push_back(static_cast<std::uint16_t>(cls[r]));
        ir.state_offset.push_back(ir.state_offset.back() + static_cast<std::uint64_t>(rank));
    }
    const std::vector<double> ones(keep.size(), 1.0);
    auto out = std::make_shared<DeviceSector>(std::move(spec), keep.data(), ones.data(),
                                              static_cast<std::int64_t>(keep.size()), rank_budget_bytes);
    out->stats_.orbits = orbits;
    out->attach_irrep(std::move(ir));
    return out;
}

void DeviceSector::attach_irrep(IrrepData ir) {
    const hipStream_t st = engine_stream();
    auto put = [&](DeviceBuffer& b, const void* h, std::size_t bytes) {
        b.reset(std::max<std::size_t>(bytes, 8));
        if (bytes) QSA_HIP(hipMemcpyAsync(b.get(), h, bytes, hipMemcpyHostToDevice, st));
    };
    put(irrep_D_, ir.D.data(), ir.D.size() * sizeof(cplx));
    put(class_C_, ir.class_C.data(), ir.class_C.size() * sizeof(cplx));
    put(class_rank_, ir.class_rank.data(), ir.class_rank.size());
    put(rep_class_, ir.rep_class.data(), ir.rep_class.size() * 2);
    put(state_offset_, ir.state_offset.data(), ir.state_offset.size() * 8);
    QSA_HIP(hipStreamSynchronize(st));
    states_ = static_cast<std::int64_t>(ir.state_offset.back());
    stats_.device_bytes += irrep_D_.bytes() + class_C_.bytes() + class_rank_.bytes() + rep_class_.bytes() +
                           state_offset_.bytes();
    irrep_ = std::move(ir);
}

dev::IrrepView DeviceSector::irrep_view() const {
    dev::IrrepView v;
    if (irrep_.d <= 1) return v;
    v.d = irrep_.d;
    v.D = irrep_D_.get<dev::Zd>();
    v.C = class_C_.get<dev::Zd>();
    v.class_rank = class_rank_.get<std::uint8_t>();
    v.rep_class = rep_class_.get<std::uint16_t>();
    v.offset = state_offset_.get<std::uint64_t>();
    return v;
}

DeviceProgram::DeviceProgram(const Program& p) {
    n_groups = static_cast<std::uint32_t>(p.n_groups());
    n_terms = p.n_terms();
    flip = upload(p.group_flip.data(), p.group_flip.size());
    setbits = upload(p.group_setbits.data(), p.group_setbits.size());
    vbegin = upload(p.group_vbegin.data(), p.group_vbegin.size());
    vval = upload(p.vsub_val.data(), p.vsub_val.size());
    vtbegin = upload(p.vsub_tbegin.data(), p.vsub_tbegin.size());
    sign = upload(p.term_sign.data(), p.term_sign.size());
    std::vector<double> re(n_terms), im(n_terms);
    real = true;
    for (std::size_t k = 0; k < n_terms; ++k) {
        re[k] = p.term_coeff[k].real();
        im[k] = p.term_coeff[k].imag();
        if (im[k] != 0.0) real = false;
    }
    cre = upload(re.data(), re.size());
    cim = upload(im.data(), im.size());
}

dev::ProgramView DeviceProgram::view() const {
    dev::ProgramView v;
    v.n_groups = n_groups;
    v.group_flip = flip.get<std::uint64_t>();
    v.group_setbits = setbits.get<dev::i32>();
    v.group_vbegin = vbegin.get<std::uint32_t>();
    v.vsub_val = vval.get<std::uint64_t>();
    v.vsub_tbegin = vtbegin.get<std::uint32_t>();
    v.term_sign = sign.get<std::uint64_t>();
    v.coeff_re = cre.get<double>();
    v.coeff_im = cim.get<double>();
    return v;
}

}  // namespace qsa
