// A symmetry sector resident in HBM (the reference's RepSectorData + its device mirror,
// include/ed/basis/rep_sector.h:134-408, src/gpu/rep_matvec.cu:253-383), built on the device.
#pragma once

#include <array>
#include <memory>
#include <vector>

#include "../device/dev_common.h"
#include "group.h"
#include "hip_util.h"

namespace qsa {

struct SectorBuildStats {
    double seconds = 0.0;
    std::int64_t orbits = 0;           ///< orbits of the subspace before the character filter
    std::uint64_t subspace_dim = 0;
    std::uint64_t device_bytes = 0;
    bool rank_table = false;           ///< the rank directory is present (else lookups binary-search)
};

/// An irrep of dimension d > 1 on a sector (the reference's rep_sector.h d > 1 fields): D(g) per
/// element; per stabiliser class its rank and d x d matrix C (C^dag M C = I on the range of
/// M = sum over the stabiliser of conj(D(s)), columns from the rank on zero); per representative its
/// class and the index of its first state. d = 1: the vectors stay empty.
struct IrrepData {
    int d = 1;
    std::vector<cplx> D;                      ///< G x d x d, row-major
    std::vector<cplx> class_C;                ///< per class: d x d, row-major
    std::vector<std::uint8_t> class_rank;     ///< per class
    std::vector<std::uint16_t> rep_class;     ///< per representative
    std::vector<std::uint64_t> state_offset;  ///< per representative, and one past the last
};

class DeviceSector {
public:
    /// The sector of an irrep of dimension d (2 .. the device kernels' limit) of spec's group:
    /// D holds D(g) for every element (G x d x d, row-major, a unitary representation); spec's
    /// characters become the traces of D (the reference's group_sector_irrep_from_table,
    /// src/engine/group_sector.cpp:78-139). Every orbit is enumerated, each representative's
    /// stabiliser read on the device, and the representatives whose class has rank 0 dropped.
    static std::shared_ptr<DeviceSector> make_irrep(GroupSpec spec, int d, const std::vector<cplx>& D,
                                                    std::uint64_t rank_budget_bytes);

    /// Enumerate the sector of `spec` on the device. `rank_budget_bytes` caps the rank directory
    /// (16 bytes per 64 states of the subspace: C(N, n_up) / 4 bytes); above it lookups
    /// binary-search the representatives.
    DeviceSector(GroupSpec spec, std::uint64_t rank_budget_bytes);
    /// Adopt representatives and 1/norms computed elsewhere (the reference's RepSectorData): the
    /// group tables and the rank directory are built here; every representative is checked to be a
    /// state of the subspace and to find itself through the group's lookup (order and canonical form).
    DeviceSector(GroupSpec spec, const std::uint64_t* reps, const double* inv_norms, std::int64_t n,
                 std::uint64_t rank_budget_bytes);

    ~DeviceSector() { fence_before_release(); }
    DeviceSector(const DeviceSector&) = delete;
    DeviceSector& operator=(const DeviceSector&) = delete;

    const GroupSpec& spec() const { return spec_; }
    /// Basis states: the representatives (d = 1), or the sum of their ranks (d > 1).
    std::int64_t dim() const { return irrep_.d > 1 ? states_ : dim_; }
    /// Representatives.
    std::int64_t n_reps() const { return dim_; }
    int irrep_dim() const { return irrep_.d; }
    const IrrepData& irrep() const { return irrep_; }
    dev::IrrepView irrep_view() const;
    const SectorBuildStats& stats() const { return stats_; }
    dev::SectorView view() const;
    bool real_characters() const { return spec_.real_characters(); }

    /// Host copies (tests, vector expansion).
    std::vector<std::uint64_t> reps() const;
    std::vector<double> inv_norms() const;

private:
    void attach_irrep(IrrepData ir);
    void setup_tables();
    /// The rank directory from the scan's kept bitmap and its exclusive scan, or (both null) from
    /// the representatives.
    void build_rank_dir(std::uint64_t total, std::uint64_t rank_budget_bytes, const std::uint64_t* kept,
                        const std::uint64_t* offsets);
    GroupSpec spec_;
    std::int64_t dim_ = 0;
    SectorBuildStats stats_;
    DeviceBuffer lut_, flips_, chi_re_, chi_im_, to_key_, cand_rec_, cand_, binom_;
    DeviceBuffer reps_, inv_norms_, rank_dir_;
    DeviceBuffer nib_;                          ///< GroupSpec::nibble_tables
    std::array<std::size_t, 3> nib_sizes_{};    ///< entries of its images, flips and key tables
    IrrepData irrep_;
    std::int64_t states_ = 0;
    DeviceBuffer irrep_D_, class_C_, class_rank_, rep_class_, state_offset_;
};

/// A compiled program in device memory.
struct DeviceProgram {
    explicit DeviceProgram(const struct Program& p);
    ~DeviceProgram() { fence_before_release(); }
    DeviceProgram(const DeviceProgram&) = delete;
    DeviceProgram& operator=(const DeviceProgram&) = delete;
    dev::ProgramView view() const;
    std::uint32_t n_groups = 0;
    std::size_t n_terms = 0;
    bool real = true;
    DeviceBuffer flip, setbits, vbegin, vval, vtbegin, sign, cre, cim;
};

}  // namespace qsa
