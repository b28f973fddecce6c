// Matrix elements of many observables between many vector pairs in one walk (the reference's
// rep_matrix_elements, include/ed/ops/program.h:130-161, src/gpu/rep_matrix_elements.cu:175-340):
// out[p][o] = <bra_p| O_o |ket_p> = sum_r sum_c conj(bra_p[r]) M_o[r][c] ket_p[c], M_o the operator's
// matrix from the ket sector to the bra sector. One wavefront per workgroup, one lane per bra
// representative, walks the merged program once: per flip group one target and one lookup in the
// ket sector, then for each observable of the group (a list the same for every lane) each lane's
// terms of its value subgroup summed to its h and its entry (the CSR build's entry) weighted into the
// pairs; the 64 lanes' contributions to a (pair, observable) cell are summed by a fixed xor tree and
// added by lane 0 to the workgroup's LDS accumulator. Every sum has a fixed order (the tree, the
// rows in grid-stride order, then the workgroups in order), so the results are the same bit for bit
// from run to run.
#include "dev_common.h"
#include "launch.h"

namespace qsa::dev {

namespace {

constexpr int kObsThreads = 64;   // one wavefront per workgroup

// The value subgroup of flip group g that acts on s, or -1 (the match of connection()).
__device__ __forceinline__ i64 subgroup_of(const ProgramView& P, u64 s, u32 g) {
    const u64 v = s & P.group_flip[g];
    const i32 sb = P.group_setbits[g];
    if (sb >= 0 && popc(v) != sb) return -1;
    u32 lo = P.group_vbegin[g], hi = P.group_vbegin[g + 1];
    const u32 stop = hi;
    while (lo < hi) {
        const u32 mid = lo + (hi - lo) / 2;
        if (P.vsub_val[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    if (lo == stop || P.vsub_val[lo] != v) return -1;
    return lo;
}

__device__ __forceinline__ double wave_total(double x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

template <bool SLC>
__global__ __launch_bounds__(kObsThreads) void k_obs_sweep(SectorView R, SectorView C, int same, ProgramView P,
                                                           const u32* term_obs, const u32* gobs_begin,
                                                           const u32* gobs, int o0, int nob, int n_pairs,
                                                           const Zd* const* bras, const Zd* const* kets,
                                                           double* partial) {
    extern __shared__ double acc[];   // [n_pairs][nob] x (re, im)
    const int cells = n_pairs * nob;
    const int lane = static_cast<int>(threadIdx.x);
    for (int i = lane; i < 2 * cells; i += kObsThreads) acc[i] = 0.0;
    __syncthreads();
    const u32 lo = static_cast<u32>(o0), hi_o = static_cast<u32>(o0 + nob);
    // every lane runs every round (the sums over the wave need all of them); rows past the end idle
    for (i64 base = static_cast<i64>(blockIdx.x) * kObsThreads; base < R.dim;
         base += static_cast<i64>(gridDim.x) * kObsThreads) {
        const i64 r = base + lane;
        const bool row = r < R.dim;
        const u64 s = row ? R.reps[r] : 0;
        const double w = row ? R.inv_norms[r] : 0.0;
        for (u32 g = 0; g < P.n_groups; ++g) {
            // the group's observables in this chunk (the same for every lane)
            u32 q0 = gobs_begin[g], q1 = gobs_begin[g + 1];
            while (q0 < q1 && gobs[q0] < lo) ++q0;
            if (q0 == q1 || gobs[q0] >= hi_o) continue;
            const i64 vs = row ? subgroup_of(P, s, g) : -1;
            const u64 t = s ^ P.group_flip[g];
            const bool diag = same && t == s;
            i64 c = r;
            double pr = 0.0, pi = 0.0;
            bool act = vs >= 0;
            if (act && !diag) {
                c = index_and_projection<SLC>(C, t, pr, pi);
                act = c >= 0;
            }
            if (!__any(act)) continue;
            u32 k = act ? P.vsub_tbegin[vs] : 0;
            const u32 k1 = act ? P.vsub_tbegin[vs + 1] : 0;
            while (k < k1 && term_obs[k] < lo) ++k;
            for (u32 q = q0; q < q1 && gobs[q] < hi_o; ++q) {
                const u32 o = gobs[q];
                double hr = 0.0, hi = 0.0;
                bool has = false;
                for (; k < k1 && term_obs[k] == o; ++k) {
                    has = true;
                    if (popc(s & P.term_sign[k]) & 1) {
                        hr += -P.coeff_re[k];
                        hi += -P.coeff_im[k];
                    } else {
                        hr += P.coeff_re[k];
                        hi += P.coeff_im[k];
                    }
                }
                double vr = 0.0, vi = 0.0;   // the matrix entry M_o[r][c], as the CSR build writes it
                if (has) {
                    if (diag) {
                        vr = hr;
                        vi = -hi;
                    } else {
                        const double xr = hr * pr - hi * pi;
                        const double xi = hr * pi + hi * pr;
                        vr = w * xr;
                        vi = w * (-xi);
                    }
                }
                for (int p = 0; p < n_pairs; ++p) {
                    double er = 0.0, ei = 0.0;
                    if (has) {
                        const Zd b = bras[p][r], kv = kets[p][c];
                        // conj(b) v kv
                        const double mr = vr * kv.re - vi * kv.im, mi = vr * kv.im + vi * kv.re;
                        er = b.re * mr + b.im * mi;
                        ei = b.re * mi - b.im * mr;
                    }
                    er = wave_total(er);
                    ei = wave_total(ei);
                    if (lane == 0) {
                        double* cell = acc + 2 * (p * nob + static_cast<int>(o) - o0);
                        cell[0] += er;
                        cell[1] += ei;
                    }
                }
            }
        }
    }
    __syncthreads();
    double* out = partial + static_cast<u64>(blockIdx.x) * 2 * cells;
    for (int i = lane; i < 2 * cells; i += kObsThreads) out[i] = acc[i];
}

// out[p][o0 + j] = the sum over workgroups, in workgroup order, of partial[wg][p][j].
__global__ void k_obs_reduce(const double* partial, int blocks, int n_pairs, int nob, int n_obs, int o0, double* out) {
    const int cells = n_pairs * nob;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cells) return;
    double re = 0.0, im = 0.0;
    for (int b = 0; b < blocks; ++b) {
        re += partial[(static_cast<u64>(b) * cells + i) * 2];
        im += partial[(static_cast<u64>(b) * cells + i) * 2 + 1];
    }
    const int p = i / nob, j = i % nob;
    out[(static_cast<u64>(p) * n_obs + o0 + j) * 2] = re;
    out[(static_cast<u64>(p) * n_obs + o0 + j) * 2 + 1] = im;
}

}  // namespace

int obs_sweep_blocks(i64 rows) {
    const i64 b = (rows + kObsThreads - 1) / kObsThreads;
    return static_cast<int>(b < 1 ? 1 : (b > 8192 ? 8192 : b));   // 32 wavefronts per CU of 256
}

void obs_sweep(const SectorView& R, const SectorView& C, bool same, const ProgramView& P, const u32* term_obs,
               const u32* gobs_begin, const u32* gobs, int o0, int nob, int n_obs, int n_pairs,
               const Zd* const* bras, const Zd* const* kets, double* partial, double* out, hipStream_t st) {
    const int blocks = obs_sweep_blocks(R.dim);
    const std::size_t lds = static_cast<std::size_t>(n_pairs) * nob * 2 * sizeof(double);
    const bool walk = R.dim > 0 && P.n_groups > 0;
    if (walk) {
        if (C.slc_L > 0)
            hipLaunchKernelGGL(k_obs_sweep<true>, dim3(blocks), dim3(kObsThreads), lds, st, R, C, same ? 1 : 0, P,
                               term_obs, gobs_begin, gobs, o0, nob, n_pairs, bras, kets, partial);
        else
            hipLaunchKernelGGL(k_obs_sweep<false>, dim3(blocks), dim3(kObsThreads), lds, st, R, C, same ? 1 : 0, P,
                               term_obs, gobs_begin, gobs, o0, nob, n_pairs, bras, kets, partial);
    }
    const int cells = n_pairs * nob;
    hipLaunchKernelGGL(k_obs_reduce, dim3((cells + 255) / 256), dim3(256), 0, st, partial, walk ? blocks : 0, n_pairs,
                       nob, n_obs, o0, out);
}

}  // namespace qsa::dev
