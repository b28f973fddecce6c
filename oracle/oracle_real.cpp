// oracle/oracle_real.cpp -- TEST INFRASTRUCTURE ONLY (the checker; see oracle.cpp's header).
// The real-block SpMV of a merged CSR, out[r] = sum_e Re(value(e)) in[col[e]] in entry order, the
// reference's RealCsrView::spmv / ReducedSymmetryCsr::spmv_with (include/ed/matvec/reduced_csr.h:
// 93-103, 130-136): on one thread, and over `threads` OpenMP threads with a static row schedule as
// the reference's host lane runs it; the bench's CPU baseline times them.
#include <cstdint>

#include <omp.h>

extern "C" void orc_csr_matvec_real(std::int64_t rows, const std::uint64_t* row_ptr, const std::uint32_t* col,
                                    const double* val, const double* x, double* y) {
    for (std::int64_t r = 0; r < rows; ++r) {
        double acc = 0.0;
        for (std::uint64_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e) acc += val[e] * x[col[e]];
        y[r] = acc;
    }
}

extern "C" void orc_csr_matvec_real_mt(std::int64_t rows, const std::uint64_t* row_ptr, const std::uint32_t* col,
                                       const double* val, const double* x, double* y, int threads) {
#pragma omp parallel for schedule(static) num_threads(threads)
    for (std::int64_t r = 0; r < rows; ++r) {
        double acc = 0.0;
        for (std::uint64_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e) acc += val[e] * x[col[e]];
        y[r] = acc;
    }
}
