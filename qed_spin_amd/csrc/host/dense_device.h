// Whole spectra of square sector blocks on the device (rocSOLVER zheevd, strided batches per dimension).
#pragma once

#include <cstdint>
#include <vector>

#include "sector_operator.h"

namespace qsa {

/// Dense blocks from this many states go to the device under the reference's 'auto' placement
/// (kDeviceDenseMinDim, include/ed/core/device.h); smaller ones are solved on the host pool.
inline constexpr std::int64_t kDeviceDenseMinDim = 1024;
/// The largest block the device lane expands (16 GiB as a dense complex matrix).
inline constexpr std::int64_t kDenseDeviceMaxDim = 32768;

/// Every block's eigenvalues, ascending (empty for an empty block). Blocks of one dimension are
/// expanded side by side into dense column-major matrices of at most batch_bytes together and
/// solved by one batched call.
std::vector<std::vector<double>> device_spectra(const std::vector<const SectorOperator*>& ops,
                                                std::uint64_t batch_bytes = std::uint64_t{8} << 30);

}  // namespace qsa
