// oracle/oracle_real.cpp -- TEST INFRASTRUCTURE ONLY (the checker; see oracle.cpp's header).
// The real-block SpMV of a merged CSR, out[r] = sum_e Re(value(e)) in[col[e]] in entry order, the
// reference's RealCsrView::spmv / ReducedSymmetryCsr::spmv_with on one thread
// (include/ed/matvec/reduced_csr.h:93-103, 130-136); the bench's CPU baseline times it.
#include <cstdint>

extern "C" void orc_csr_matvec_real(std::int64_t rows, const std::uint64_t* row_ptr, const std::uint32_t* col,
                                    const double* val, const double* x, double* y) {
    for (std::int64_t r = 0; r < rows; ++r) {
        double acc = 0.0;
        for (std::uint64_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e) acc += val[e] * x[col[e]];
        y[r] = acc;
    }
}
