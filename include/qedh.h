/*
 * qedh.h -- the C-ABI of the MI355X exact-diagonalisation backend (libqedspin.so).
 *
 * Plain pointers and sizes, opaque handles, int status codes: 0 ok, 1 invalid request,
 * 2 resource limit (out of device memory / over a budget), 3 no device, 4 unsupported on the
 * device, 5 runtime error, 6 convergence failure; qedh_last_error() holds the message of the
 * last failure on the calling thread. The reference's error classes map one to one
 * (include/ed/core/errors.h:24-50: InvalidRequest, ResourceLimit, DeviceUnavailable,
 * DeviceUnsupported, ConvergenceError).
 *
 * Vectors passed to the apply entry points are DEVICE pointers (qedh_malloc or any HIP
 * allocation); a sector operator's vectors are double when qedh_matvec_info reports it real,
 * else interleaved complex double (std::complex<double> / hipDoubleComplex layout).
 *
 * Each entry point names the reference interface it replaces (file:line in the reference
 * checkout).
 */
#ifndef QEDH_H
#define QEDH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qedh_operator_s* qedh_operator;
typedef struct qedh_program_s* qedh_program;
typedef struct qedh_sector_s* qedh_sector;
typedef struct qedh_matvec_s* qedh_matvec;

/* ---- runtime ------------------------------------------------------------------------------ */
/* The last failure's message on this thread (empty when none). */
const char* qedh_last_error(void);
/* Visible HIP devices; 0 without a device (replaces ed::cuda_device_count / has_cuda_build,
 * python/qed/_verbs/_device.py:18-29). */
int qedh_device_count(int* count);
/* Free and total bytes of the current device (hipMemGetInfo): the free-memory probe of the
 * reference's placement, DeviceProbe::free_bytes = ed::core::available_device_bytes
 * (include/ed/core/select_backend.h:79-83, 112-160); DeviceUnavailable without a device. */
int qedh_device_memory(uint64_t* free_bytes, uint64_t* total_bytes);
/* Device memory for callers without their own allocator (replaces CudaBackend::allocate /
 * deallocate / copy_from_host / copy_to_host, include/ed/gpu/cuda_backend.cuh:104-230). */
int qedh_malloc(void** ptr, size_t bytes);
int qedh_free(void* ptr);
int qedh_memcpy_htod(void* dst, const void* src, size_t bytes);
int qedh_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int qedh_memcpy_dtod(void* dst, const void* src, size_t bytes);
/* out[0] + i out[1] = x^H y of two device vectors of n entries (complex or real), reduced on the
 * device in a fixed order (BasicBackend::dot, include/ed/matvec/backend.h:56-179). */
int qedh_vdot(const void* x, const void* y, int64_t n, int complex_vectors, double* out);
int qedh_memset(void* dst, int value, size_t bytes);
int qedh_synchronize(void);
/* The calling thread's stream for every engine launch, copy and allocation: own != 0 a non-blocking
 * stream of the engine's pool (taken on first use, handed back drained -- never destroyed -- when
 * the thread ends), 0 the null stream. Calls that return results synchronise their stream first,
 * so objects built on one thread's stream are complete when the call returns; an object dropped on
 * any thread frees its device memory only after the work every engine stream has enqueued so far.
 * Lets host threads build sectors and operators concurrently (the reference keeps every thread on
 * the legacy default stream, include/ed/matvec/batcher.h:21-23; no counterpart). */
int qedh_thread_stream(int own);
/* Cross-stream fences taken so far (one per shared object released) and streams in the pool
 * (no counterpart; tests of the release rule). Either pointer may be NULL. */
int qedh_stream_stats(uint64_t* fences, int64_t* pooled_streams);

/* ---- the Krylov drivers' vector backend on device pointers (BasicBackend<Scalar>,
 * include/ed/matvec/backend.h:56-179; its device implementation CudaBackend,
 * include/ed/gpu/cuda_backend.cuh:104-415). complex_vectors != 0: Scalar = complex<double>
 * (interleaved re, im; scalars passed as double[2]), else double (scalars double[1]). Each call
 * runs on the calling thread's engine stream and returns when done; reductions are deterministic
 * (fixed grid and order). ------------------------------------------------------------------- */
/* BasicBackend::fill_zero (backend.h:73). */
int qedh_fill_zero(int complex_vectors, void* x, size_t n);
/* y <- alpha x + y (BasicBackend::axpy, backend.h:89). */
int qedh_axpy(int complex_vectors, const double* alpha, const void* x, void* y, size_t n);
/* x <- alpha x (BasicBackend::scale, backend.h:90). */
int qedh_scale(int complex_vectors, const double* alpha, void* x, size_t n);
/* y <- alpha x + beta y (BasicBackend::axpby, backend.h:96). */
int qedh_axpby(int complex_vectors, const double* alpha, const void* x, const double* beta, void* y, size_t n);
/* out = x^H y, conjugate on the left (BasicBackend::dot, backend.h:91). */
int qedh_dot(int complex_vectors, const void* x, const void* y, size_t n, double* out);
/* out = ||x||_2 (BasicBackend::nrm2, backend.h:92). */
int qedh_nrm2(int complex_vectors, const void* x, size_t n, double* out);
/* y <- y + alpha x, then out = z^H y, one pass (BasicBackend::axpy_dot, backend.h:108-112). */
int qedh_axpy_dot(int complex_vectors, const double* alpha, const void* x, void* y, const void* z, size_t n,
                  double* out);
/* y <- y + alpha x, then out = ||y||_2, one pass (BasicBackend::axpy_nrm2, backend.h:113-116). */
int qedh_axpy_nrm2(int complex_vectors, const double* alpha, const void* x, void* y, size_t n, double* out);
/* coeffs_out[j] = basis[j]^H v for j < num_basis; basis: a HOST array of DEVICE vectors, coeffs_out
 * a host array of num_basis scalars; one pass over v feeds every column (BasicBackend::dot_many,
 * backend.h:133-137). */
int qedh_dot_many(int complex_vectors, const void* const* basis, size_t num_basis, const void* v, size_t n,
                  double* coeffs_out);
/* v += sum_j alphas[j] basis[j] (host alphas, host array of device vectors; BasicBackend::axpy_many,
 * backend.h:141-145). */
int qedh_axpy_many(int complex_vectors, const double* alphas, const void* const* basis, size_t num_basis, void* v,
                   size_t n);
/* qedh_dot_many without any allocation. No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). Same arguments and results as qedh_dot_many (bit for bit: the same kernels);
 * its column table, coefficients and partial sums live in a fixed device array the library holds on
 * each device until the process ends (about 4.3 MB). num_basis above QEDH_NOALLOC_MAX_BASIS is
 * InvalidRequest. Calls on one device share that array: calls that overlap on one device race. */
#define QEDH_NOALLOC_MAX_BASIS 4096
int qedh_dot_many_noalloc(int complex_vectors, const void* const* basis, size_t num_basis, const void* v, size_t n,
                          double* coeffs_out);
/* C = alpha op(A) op(B) + beta C, column-major device matrices, op 'N' / 'T' / 'C' (BasicBackend::gemm,
 * backend.h:156-161): the library GEMM (rocBLAS zgemm / dgemm) on the engine stream. alpha / beta NULL:
 * 1 / 0. */
int qedh_gemm(int complex_vectors, char op_a, char op_b, size_t m, size_t n, size_t k, const double* alpha,
              const void* A, size_t lda, const void* B, size_t ldb, const double* beta, void* C, size_t ldc);

/* ---- operators: canonical spin-1/2 terms (src/ops/algebra.cpp, include/ed/ops/algebra.h) ----- */
int qedh_operator_create(int n_sites, qedh_operator* out);
/* coeff * ops[0](sites[0]) ops[1](sites[1]) ..., the last factor acting first; ops from
 * "+-zxyudI" (MaskedOperator::product, src/ops/algebra.cpp:67-100). */
int qedh_operator_add_product(qedh_operator op, const char* ops, const int* sites, double re, double im);
/* s_H (sum |c|), number of canonical terms, realness (Operator::norm_bound, isReal). */
int qedh_operator_info(qedh_operator op, int64_t* n_terms, double* norm_bound, int* is_real);
int qedh_operator_destroy(qedh_operator op);

/* ---- programs (src/ops/program.cpp:34-141, include/ed/ops/operator.h:132-138) ---------------- */
/* The flip-grouped program of op (adjoint != 0: of op^dagger, the row program of a sector walk). */
int qedh_program_compile(qedh_operator op, int adjoint, qedh_program* out);
/* The lambda-projected program of op (adjoint: of op^dagger) with ket sector src and bra sector
 * tgt of one group (compile_program, src/ops/program.cpp:61-130). */
int qedh_program_compile_projected(qedh_operator op, int adjoint, qedh_sector src, qedh_sector tgt,
                                   qedh_program* out);
int qedh_program_counts(qedh_program p, int64_t* n_groups, int64_t* n_vsub, int64_t* n_terms);
int qedh_program_export(qedh_program p, uint64_t* group_flip, int32_t* group_setbits, uint32_t* group_vbegin,
                        uint64_t* vsub_val, uint32_t* vsub_tbegin, uint64_t* term_sign, double* term_coeff);
/* A program from the reference's MaskedProgram arrays as they are (group_vbegin[n_groups + 1],
 * vsub_tbegin[n_vsub + 1], term_coeff[2 n_terms] re/im): the adapter's way in for a program the
 * reference already compiled (include/ed/ops/program.h:66-95). Offsets are checked. */
int qedh_program_import(int64_t n_groups, const uint64_t* group_flip, const int32_t* group_setbits,
                        const uint32_t* group_vbegin, int64_t n_vsub, const uint64_t* vsub_val,
                        const uint32_t* vsub_tbegin, int64_t n_terms, const uint64_t* term_sign,
                        const double* term_coeff, qedh_program* out);
int qedh_program_destroy(qedh_program p);

/* ---- sectors: enumerated on the device (include/ed/basis/orbit_table.h:250-286,
 * src/engine/group_sector.cpp:52-76 + 159-205, include/ed/basis/rep_sector.h:134-408) ----------- */
/* perms: group_size x n_sites row-major (bit i of g s = bit perms[g n + i] of s); flips: per
 * element XOR mask or NULL; chars: group_size complex characters (re, im); n_up < 0: the full
 * 2^N space. sublattice_mode: -1 the size rule (relaxed != 0: 16 distinct permutations, else 64),
 * 0 never, 1 whenever a block system exists (ED_SYM_SUBLATTICE,
 * include/ed/basis/sublattice_code.h:187-230). rank_budget_bytes caps the rank directory (16 bytes per 64 states of the
 * subspace); above it lookups binary-search the representatives. */
int qedh_sector_create(int n_sites, int n_up, int group_size, const int* perms, const uint64_t* flips,
                       const double* chars, int sublattice_mode, int relaxed, uint64_t rank_budget_bytes,
                       qedh_sector* out);
/* A sector from representatives and 1/norms computed elsewhere -- the reference's RepSectorData
 * (reps, inv_norms of include/ed/basis/rep_sector.h:134-408) handed over as they are: the group's
 * tables (LUT, sublattice code per sublattice_mode / relaxed as in qedh_sector_create) and the rank
 * table are built on the device, and every representative is checked to be a state of the subspace
 * whose lookup finds itself (plain or key order, as the group's code has it); InvalidRequest names
 * the first that does not. The 1/norms are taken as given. */
int qedh_sector_import(int n_sites, int n_up, int group_size, const int* perms, const uint64_t* flips,
                       const double* chars, int sublattice_mode, int relaxed, int64_t n_reps, const uint64_t* reps,
                       const double* inv_norms, uint64_t rank_budget_bytes, qedh_sector* out);
/* The sublattice key-order tables of a group without building a sector (host only, no device):
 * the reference's SublatticeCode::of (include/ed/basis/sublattice_code.h:187-230). *block gets L
 * (0: not engaged; nothing else is written). key_bit[n_sites]; lead[(n_sites / L) 2^L] the least
 * leading-block pattern reachable from each block pattern; cand_off[(n_sites / L) 2^L + 1] and
 * cand[*n_cand] the elements reaching it, ascending. lead / cand_off / cand may be NULL (sizes
 * only); a cand_cap below the count is InvalidRequest. */
int qedh_sublattice_tables(int n_sites, int group_size, const int* perms, const uint64_t* flips, int sublattice_mode,
                           int relaxed, int* block, int* key_bit, uint16_t* lead, uint32_t* cand_off, uint16_t* cand,
                           int64_t cand_cap, int64_t* n_cand);
/* The sector of an irrep of dimension irrep_dim of the group (the reference's
 * group_sector_irrep_from_table, src/engine/group_sector.cpp:78-139, with the d > 1 fields of
 * include/ed/basis/rep_sector.h:198-215): D holds D(g) for every element, group_size x d x d complex
 * (re, im) row-major, a unitary representation. Every orbit is enumerated (least member, plain or
 * key order), each representative's stabiliser S read on the device; per stabiliser class C spans
 * the range of sum_{s in S} conj(D(s)) (rank(C) states per representative, C^dag M C = I), and the
 * representatives of rank 0 are dropped. The sector's basis states are the partner-0 states
 * |r; a> = sqrt(d/|G|) sum_j C_r[j][a] sum_g conj(D(g)[0][j]) |g r>, a < rank(r); its dim counts
 * them. irrep_dim 1: D holds the characters (qedh_sector_create). d above 8 (the reference's
 * kMaxIrrepDim, include/ed/matvec/sector_rows.h:66): DeviceUnsupported. */
int qedh_sector_create_irrep(int n_sites, int n_up, int group_size, const int* perms, const uint64_t* flips,
                             int irrep_dim, const double* D, int sublattice_mode, int relaxed,
                             uint64_t rank_budget_bytes, qedh_sector* out);
/* d (1 for a sector of characters), the representatives, the stabiliser classes (0 for d = 1). */
int qedh_sector_irrep_info(qedh_sector s, int* irrep_dim, int64_t* n_reps, int64_t* n_classes);
/* A d > 1 sector's fields: rep_class[n_reps], state_offset[n_reps + 1] (the index of each
 * representative's first state), class_rank[n_classes], class_C[n_classes][d][d] complex (re, im)
 * row-major (columns from the rank on zero); each may be NULL. d = 1: InvalidRequest. */
int qedh_sector_irrep_export(qedh_sector s, uint16_t* rep_class, uint64_t* state_offset, uint8_t* class_rank,
                             double* class_C);
/* dim (basis states), orbits before the character filter, sublattice block size (0: plain order),
 * build seconds. */
int qedh_sector_info(qedh_sector s, int64_t* dim, int64_t* n_orbits, int* sublattice_block, double* build_seconds);
/* Matrix elements of many observables between many vector pairs in one walk of the bra sector
 * (rep_matrix_elements, include/ed/ops/program.h:130-161, and its device sweep
 * src/gpu/rep_matrix_elements.cu:175-340): out[p][o] (complex, re/im) = <bras[p]| O_o |kets[p]>,
 * programs[o] the program of O_o from the ket sector to the bra sector as qedh_matvec_create takes it
 * (qedh_program_compile_projected of O_o^dag with the bra sector as src, or qedh_program_compile for
 * one sector). bras / kets: host arrays of n_pairs device vectors (complex, of each sector's dim);
 * 1 <= n_pairs <= 4096. The observables' terms are merged into one program, so a flip group costs
 * one target lookup for all of them; every sum has a fixed order (a wavefront's lanes by a fixed
 * tree, rows in order, workgroups in order), so a call's results are the same bit for bit from run to
 * run and whatever other pairs share the call; against other evaluations they agree to rounding
 * (~1e-13 relative). Sectors of irreps of dimension > 1: DeviceUnsupported. */
int qedh_matrix_elements(qedh_sector bra_sector, qedh_sector ket_sector, int n_obs, const qedh_program* programs,
                         int n_pairs, const void* const* bras, const void* const* kets, double* out);
/* The representatives (n_reps of qedh_sector_irrep_info; dim for d = 1) and their 1/norms (1 for a
 * d > 1 sector, whose normalisation is in C). */
int qedh_sector_export(qedh_sector s, uint64_t* reps, double* inv_norms);
/* key_bit[n_sites] of the sublattice order; returns the block size through *block (0: none). */
int qedh_sector_key_bits(qedh_sector s, int* key_bit, int* block);
/* index_and_projection of n host states on the device (rep_symmetry_basis_policy.h:221-241). */
int qedh_sector_index_and_projection(qedh_sector s, const uint64_t* states, int64_t n, int64_t* index,
                                     double* proj);
/* Sector vectors expanded into a basis without symmetry (ed::sectors::expand, src/engine/eigs.cpp:699-747,
 * behind EigResult.vectors(basis=...), python/qed/_verbs/eigs.py:52-78): u is a DEVICE buffer of nvec
 * vectors of dim complex entries; out a DEVICE buffer of nvec x out_dim complex entries, out_dim =
 * C(N, n_up) (full = 0: the Sz sector, states in ascending order = colex rank) or 2^N (full != 0,
 * N <= 34). Each vector is normalised to unit length; norms[nvec] (or NULL) gets the norms before.
 * Refused: a sector without fixed n_up into the Sz basis, an image outside the Sz sector, a vector
 * the expansion annihilates. */
int qedh_sector_expand(qedh_sector s, const void* u, int nvec, int full, void* out, double* norms);
int qedh_sector_destroy(qedh_sector s);

/* ---- sector operators (RepSectorMatVec / CrossSectorMatVec device lanes,
 * src/engine/internal.h:163-609; make_sector_matvec_gpu_rep, make_cross_matvec_gpu_rep,
 * build_sector_csr_gpu, csr_matvec_gpu(_multi), include/ed/gpu/rep_matvec.h:43-101) ----------- */
enum { QEDH_LANE_AUTO = 0, QEDH_LANE_GATHER = 1, QEDH_LANE_CSR = 2 };
/* OR-ed into the lane of qedh_matvec_create (no counterpart in the reference): skip the SpMV autotuning
 * of a large CSR operator (for operators whose applies run through a batched engine instead). */
enum { QEDH_LANE_NO_TUNE = 0x100 };
typedef struct {
    int64_t rows, cols;
    int is_real;          /* vectors are double */
    int lane;             /* QEDH_LANE_GATHER or QEDH_LANE_CSR */
    int64_t nnz;          /* CSR entries (merged), 0 on the gather lane */
    int value_bytes;      /* 1 / 2 (dictionary ids) or 8 / 16 (full values) */
    int id_width_ref;     /* the reference's rule: 1, 2 or 16 */
    int64_t dict_size;
    uint64_t csr_bytes;
    double build_seconds;
    int lanes;            /* SpMV lanes per row */
    double norm_bound;
} qedh_matvec_info_t;
/* cols == NULL or == rows: one sector (the diagonal shortcut applies). prog: the row program
 * (the walk of O^dagger; for two sectors the projected program of O^dagger with rows as ket).
 * Sectors of an irrep of dimension d > 1 (CSR lane only): two of them must carry the same D, and
 * prog is the plain walk of O^dagger for an O that commutes with the group (total S+-, S^2; the
 * reference's CrossSectorMatVec, src/engine/internal.h:557-563); else InvalidRequest. */
int qedh_matvec_create(qedh_sector rows, qedh_sector cols, qedh_program prog, double norm_bound, int lane,
                       uint64_t csr_budget_bytes, qedh_matvec* out);
int qedh_matvec_info(qedh_matvec m, qedh_matvec_info_t* info);
/* out = O in (bind_cuda, include/ed/matvec/linear_operator.h:51-54); complex_vectors != 0 applies a
 * real operator to complex vectors. */
int qedh_matvec_apply(qedh_matvec m, const void* in, void* out, int complex_vectors);
/* The same on host buffers (the reference's LinearOperator::apply(const Complex* in, Complex* out,
 * size_t n), include/ed/matvec/linear_operator.h:32-77): in copied to the device, applied, out copied
 * back; real doubles when the operator is real and complex_vectors is 0, else complex pairs. */
int qedh_matvec_apply_host(qedh_matvec m, const void* in, void* out, int complex_vectors);
/* outs[i] = O ins[i] (bind_cuda_multi; k-vector applies equal single applies bit for bit). */
int qedh_matvec_apply_multi(qedh_matvec m, const void* const* ins, void* const* outs, int k, int complex_vectors);
/* Block-interleaved X[row * nv + i], nv in {1, 2, 4, 8}; CSR lane. */
int qedh_matvec_apply_block(qedh_matvec m, int nv, const void* X, void* Y, int complex_vectors);
int qedh_matvec_set_lanes(qedh_matvec m, int lanes);
/* SpMV tuning (no counterpart in the reference). Large tuned operators also keep a SELL-64 copy of
 * the CSR (64-row slices, one lane per row, sums in entry order) and apply through it; the lanes
 * per row then do not apply. Lanes per row (0 keeps the current) and flags:
 * non-temporal streaming of the CSR arrays, one contiguous row range per XCD, the single-vector
 * gathers issued two, four or eight entries at a time, u8-id dictionaries read from global memory
 * instead of LDS, 256 or 1024 threads per workgroup instead of 512 (every flag leaves the results
 * bit for bit unchanged). The default is XCD rows, unroll 4 (or what the autotuner found), LDS
 * dictionary, 512 threads; flags replace it. */
enum {
    QEDH_TUNE_NONTEMPORAL = 1,
    QEDH_TUNE_XCD_ROWS = 2,
    QEDH_TUNE_UNROLL = 4,
    QEDH_TUNE_UNROLL8 = 8,
    QEDH_TUNE_GLOBAL_DICT = 16,
    QEDH_TUNE_WG256 = 32,
    QEDH_TUNE_WG1024 = 64,
    QEDH_TUNE_UNROLL2 = 128,
    QEDH_TUNE_UNROLL16 = 256,   /* SELL-64 layout: sixteen gathers in flight per lane */
    QEDH_TUNE_CSR = 512,        /* apply through the CSR layout even where a SELL-64 copy exists */
    QEDH_TUNE_SELL = 1024,      /* build the SELL-64 copy now, whatever the operator's size */
    QEDH_TUNE_SELL_SUMS1 = 2048, /* SELL-64: one sum per row instead of four in a fixed tree (other bits) */
    QEDH_TUNE_SELL_DIAG = 4096  /* SELL-64 rebuilt with each row's diagonal entry first (other bits) */
};
int qedh_matvec_set_tuning(qedh_matvec m, int lanes, int flags);
/* The merged CSR (download_csr): row_ptr[rows+1], col[nnz], val[2 nnz] (re, im). */
int qedh_matvec_export_csr(qedh_matvec m, uint64_t* row_ptr, uint32_t* col, double* val);
/* Rows [r0, r1) of it: row_ptr[r1 - r0 + 1] (absolute offsets), then col / val of those rows
 * (pass NULL col / val to read the offsets first). */
int qedh_matvec_export_csr_rows(qedh_matvec m, int64_t r0, int64_t r1, uint64_t* row_ptr, uint32_t* col,
                                double* val);
int qedh_matvec_destroy(qedh_matvec m);
/* Average seconds of one apply over `iters` timed applies (after `warmup`) of k vectors at once
 * (k = 1: the single apply; 2..8 multi-vector; block != 0: block-interleaved), measured with HIP
 * events on the stream the kernels run on. Vectors are allocated inside and filled with random unit
 * vectors (Gaussian entries, normalised; the block form normalises the whole block). */
int qedh_matvec_benchmark(qedh_matvec m, int k, int block, int complex_vectors, int warmup, int iters,
                          double* avg_seconds);

/* ---- Krylov drivers (include/ed/krylov/lanczos.h:266-576, src/engine/block_solve.cpp:433-655,
 * include/ed/thermal/ftlm.h:277-480, include/ed/dynamics/cf.h:103-161) -------------------------- */
/* One Lanczos run from a host start vector (the operator's scalar type unless complex_vectors):
 * reorth 0 None, 1 FullCGS2, 2 LocalDGKS3. alpha[max_iter], beta[max_iter + 1]; *iters steps
 * done; timings[4] = apply s, reorth s, other vector s, total s (may be NULL). */
int qedh_lanczos(qedh_matvec m, const void* v0_host, int complex_vectors, int max_iter, int reorth,
                 double breakdown_tol, double* alpha, double* beta, int* iters, double* timings);
/* The lowest k levels of the block (solve_block_lowest): the k = 1 scan or Krylov-Schur.
 * values[k]; *n_values found; *converged; *applies. */
int qedh_block_lowest(qedh_matvec m, int k, int64_t max_iter, double* values, int* n_values, int* converged,
                      int64_t* applies);
/* The pruning estimate of a block (estimate_lowest, src/engine/block_solve.cpp:630-655, as the eigs
 * walk prunes with it, src/engine/eigs.cpp:129-155, 450-560): 40 Lanczos steps without
 * reorthogonalisation from the fixed start seeded 0xE57A7E; *theta the lowest Ritz value (an upper
 * bound on the block's lowest level), *residual its residual bound; -inf when nothing ran. */
int qedh_block_estimate(qedh_matvec m, double* theta, double* residual, int64_t* applies);
/* The lowest k eigenpairs of the block (Krylov-Schur with locking at 1e-10 s_H, the vectors of
 * solve_block_eigenpairs, src/engine/block_solve.cpp:600-655): values[k] ascending, vectors: a
 * DEVICE buffer of k x dim complex entries (vector i at vectors + i dim), normalised. */
int qedh_block_eigenpairs(qedh_matvec m, int k, int64_t max_iter, double* values, void* vectors, int* n_values,
                          int* converged);
/* FTLM of the block: samples from splitmix64(seed + phi (s + 1)); lnZ/E/V per beta; *e_min. */
int qedh_ftlm(qedh_matvec m, int samples, int krylov, uint64_t seed, int n_beta, const double* betas,
              double breakdown_tol, double* lnZ, double* E, double* V, double* e_min);
/* The continued fraction from a host vector phi (complex): S[n_omega]; has_shift selects
 * energy_shift, else the lowest Ritz value; out_shift, out_norm_sq, out_change may be NULL. */
int qedh_cf(qedh_matvec m, const double* phi_host, int krylov, int n_omega, const double* omega, double eta,
            int has_shift, double energy_shift, double breakdown_tol, double* S, double* out_shift,
            double* out_norm_sq, double* out_change);

/* ---- Batched Krylov engine: many blocks x R vectors per block in one packed multi-vector
 * matvec per step, the step captured as a hipGraph (replaces the per-block, lockstep-of-8 sample
 * loop of include/ed/thermal/ftlm.h:277-480 driven by src/engine/thermal.cpp:161-344, 471-535) ---- */
/* Lanczos without a kept basis of R vectors in each of n_blocks square CSR-lane blocks: reorth 2
 * LocalDGKS3 with K = 1 (FTLM), 0 none (the continued fraction).
 * start: the complex start vectors in batch layout, block b vector s at entry (off_b + i) R + s,
 * off_b = the summed dimensions of the blocks before b (normalised inside). breakdown_tol[b] per
 * block. Outputs alpha[krylov][n_blocks][R], beta[krylov][n_blocks][R] (beta of step j is
 * beta_{j+1}), steps[n_blocks][R]. timings[3] = pack s, steps s, total s (may be NULL). */
int qedh_batch_lanczos(const qedh_matvec* blocks, int n_blocks, int R, int krylov, const double* start,
                       const double* breakdown_tol, int reorth, int use_graph, double* alpha, double* beta,
                       int* steps, double* timings);
/* FTLM of n_blocks blocks at once: block b samples s = 0..samples-1 from
 * splitmix64(block_seeds[b] + phi (s + 1)) as qedh_ftlm does, breakdown_tol[b] per block (NULL:
 * 64 eps times the block's norm bound, kBreakdownRel of include/ed/core/numerics.h:23).
 * lnZ/E/V[n_blocks][n_beta], e_min[n_blocks]. stats[10] (may be NULL) = pack s, seed s, steps s,
 * post s, total s, block-vector applies, packed rows, packed nnz, the tile schedule the engine's
 * autotuner chose (0 in order, 1 grouped by XCD; results are the same bit for bit under either),
 * its tuning s. QED_SPIN_AMD_BATCH_SCHEDULE=0|1 pins the schedule. */
int qedh_ftlm_batched(const qedh_matvec* blocks, int n_blocks, int samples, int krylov, const uint64_t* block_seeds,
                      const double* breakdown_tol, int n_beta, const double* betas, int use_graph, double* lnZ,
                      double* E, double* V, double* e_min, double* stats);
/* Continued fractions of many (block, vector) pairs at once (cf_spectral_from_vector,
 * include/ed/dynamics/cf.h:103-161, for every probe image the dynamics verb forms,
 * src/engine/dynamics.cpp:528-560): starts[b R + s] a DEVICE complex vector of block b's dimension
 * (NULL: no pair); Lanczos without reorthogonalisation, krylov steps, breakdown_tol (<= 0: 1e-12);
 * shift = energy_shift when has_shift, else each pair's lowest Ritz value. Outputs
 * S[n_blocks R][n_omega] (weighted by ||phi||^2), norm_sq, shifts, steps (each may be NULL but S),
 * stats[9] = pack s, steps s, host s, total s, applies, rows, nnz, tile schedule, tuning s (may be
 * NULL). */
int qedh_cf_batched(const qedh_matvec* blocks, int n_blocks, int R, const void* const* starts, int krylov,
                    int has_shift, double energy_shift, double breakdown_tol, int n_omega, const double* omega,
                    double eta, int use_graph, double* S, double* norm_sq, double* shifts, int* steps, double* stats);
/* Finite-temperature dynamics of a whole walk of source sectors (ftlm_dynamics_kernel,
 * include/ed/dynamics/ftlm_dynamics.h:110-397, for every source of the T > 0 dynamics verb,
 * src/engine/dynamics.cpp:693-1257; autocorrelations A = B). Source b's samples are drawn from
 * splitmix64(source_seeds[b] + phi (s + 1)) (std::mt19937 -> std::normal_distribution, on the
 * device); a channel c is a probe channel_probe[c] (CSR lane) from source channel_source[c] to a
 * target sector whose H is channel_h[c]. All sources' samples and all channels' target runs advance
 * through the packed multi-vector engine with kept bases (full CGS2 on sectors of at most
 * max(4 krylov, 2^16) states, LocalDGKS3 on larger ones); the overlaps <A v_a|w_b> are products of
 * the kept bases split over row chunks, summed per pair in a fixed order; waves of sources keep the
 * kept vectors within budget_bytes (0: two fifths of the device's memory, at most nine tenths of what
 * is free at the call). Outputs,
 * un-normalised (times each source's dimension / samples done): Z[n_sources][n_t], S[n_channels][n_t][n_omega] complex
 * (re, im); e_min[n_sources] each source's samples' common reference and samples_done[n_sources]
 * (each may be NULL); stats[6] = Lanczos s, overlap s, host s, total s, waves, set-up s (the wave plan and
 * the kept-vector regions; or NULL). */
int qedh_ftlm_dynamics_walk(int n_sources, const qedh_matvec* source_h, const uint64_t* source_seeds,
                            const double* breakdown_tol, int n_channels, const int* channel_source,
                            const qedh_matvec* channel_h, const qedh_matvec* channel_probe, int samples, int krylov,
                            int n_t, const double* temperatures, int n_omega, const double* omega, double eta,
                            uint64_t budget_bytes, double* Z, double* S, double* e_min, int* samples_done,
                            double* stats);
/* The same for one source sector: qedh_ftlm_dynamics_walk with n_sources = 1, the targets as its
 * channels, seed its block seed; Z[n_t], S[n_targets][n_t][n_omega], *e_min, *samples_done. */
int qedh_ftlm_dynamics(qedh_matvec source_h, int n_targets, const qedh_matvec* target_h, const qedh_matvec* probes,
                       int samples, int krylov, uint64_t seed, double breakdown_tol, int n_t,
                       const double* temperatures, int n_omega, const double* omega, double eta, double* Z,
                       double* S, double* e_min, int* samples_done);
/* The FTLM start vectors of samples first_sample .. first_sample + count - 1 of seed `base`, drawn
 * on the device: each std::mt19937(low 32 bits of splitmix64(base + phi (s + 1))) ->
 * std::normal_distribution<double>(0, 1), entry i = (n_2i, n_2i+1), not normalised (the reference's
 * sample_engine and gaussian_vector, include/ed/thermal/sample_seed.h:37-42, 83-89, as ftlm.h:316
 * and ftlm_dynamics.h:157 draw them). out: device, [count][n] complex (re, im). The entries equal the
 * host's up to the last bit of the logarithm. */
int qedh_gaussian_vectors(uint64_t base, uint64_t first_sample, int count, int64_t n, double* out);
/* The whole spectrum of blocks, ascending, by a dense solve (the reference's dense lanes of the eigs
 * and spectrum walks: solve_block_dense, src/engine/block_solve.cpp:164-262, and the DenseBatch lane,
 * src/gpu/little_group.cu:38-211): blocks of at least device_min_dim states (< 0: 1024, the
 * reference's kDeviceDenseMinDim under 'auto'; 0: every block) and every block above 8192 on the
 * device (CSR expanded to a dense matrix on the device, rocSOLVER zheevd in strided batches per
 * dimension), the others on the host (Householder tridiagonalisation and implicit QL, blocks in
 * parallel). values: the blocks' spectra one after another (dim of each block entries); blocks of at
 * most 32768 states. */
int qedh_block_spectrum(const qedh_matvec* blocks, int n_blocks, int64_t device_min_dim, double* values);
/* Exact thermodynamics of blocks solved densely (the reference's dense lanes for blocks of at most
 * dense_max_dim states, src/engine/thermal.cpp:197-200, curves of include/ed/thermal/curves.h:38-65):
 * each block's whole spectrum as qedh_block_spectrum solves it (blocks from 1024 states on the
 * device), then its curves. lnZ/E/V[n_blocks][n_beta], e_min[n_blocks]; blocks of at most 32768
 * states. */
int qedh_exact_thermal(const qedh_matvec* blocks, int n_blocks, int n_beta, const double* betas, double* lnZ,
                       double* E, double* V, double* e_min);
/* The system's curves from its blocks' (src/engine/thermal.cpp:597-640): weights[n_blocks]
 * (multiplicities), inputs [n_blocks][n_beta]; outputs [n_beta]: lnZ, E, C = beta^2 Var, S, F. */
int qedh_thermal_combine(int n_blocks, int n_beta, const double* betas, const double* weights, const double* lnZ,
                         const double* E, const double* V, double* out_lnZ, double* out_E, double* out_C,
                         double* out_S, double* out_F);

/* ---- Serving (no counterpart in the reference, which has no serving layer): a stream of solves
 * on spin chains, admitted against an HBM budget and run concurrently on slots with a HIP stream
 * each, sectors kept resident between requests ------------------------------------------------ */
/* One request: the periodic chain of n_sites spins with J1 on (i, i+1), J2 on (i, i+2) and
 * chi S_i . (S_{i+1} x S_{i+2}), in the sector n_up = n_sites / 2, momentum 0. kind 0: the k lowest
 * levels; 1: FTLM (samples, depth, seed) at the options' temperatures; 2: S^z(q) for q the lattice
 * momentum nearest pi from the ground state (continued fraction of depth krylov). */
typedef struct {
    int kind;
    int n_sites;
    double j1, j2, chi;
    double arrival_s;
    int k, samples, depth, krylov;
    uint64_t seed;
} qedh_request_t;
typedef struct {
    int slots;          /* requests in flight */
    int batched;        /* FTLM samples through one multi-vector apply (else one after another) */
    int cache_sectors;  /* keep sectors resident between requests */
    int realtime;       /* honour arrival times (else all requests are there at 0) */
    uint64_t hbm_budget, cache_budget;   /* 0: defaults (200 GiB, 48 GiB) */
    int n_beta;
    const double* betas;
    int n_omega;
    const double* omega;
    double eta;
    double progress_s;  /* > 0: a progress line on stderr at most this often */
} qedh_serve_options_t;
/* Run the stream. values[n][stride]: eigs levels / thermal lnZ, E, V per beta / dynamics S(omega);
 * e0[n]; timing[n][4] = arrival, start, finish, latency (s); status[n] (0 ok); dims[n];
 * stats[16] = wall s, solves/s, p50, p99, mean, max latency, completed, failed, peak reserved bytes,
 * peak in flight, sector builds, sector cache hits, seconds of the requests spent building sectors,
 * programs and CSRs (summed), and start-to-finish seconds summed per kind (eigs, thermal, dynamics).
 * Every output but values may be NULL. */
int qedh_serve(const qedh_request_t* reqs, int n, const qedh_serve_options_t* opt, int stride, double* values,
               double* e0, double* timing, int* status, int64_t* dims, double* stats);
/* The HBM bytes a request reserves at admission. */
int qedh_request_footprint(const qedh_request_t* req, int batched, uint64_t* bytes);

#ifdef __cplusplus
}
#endif

#endif /* QEDH_H */
