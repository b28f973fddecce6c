// Matrix elements of many observables between many vector pairs in one walk (the reference's
// rep_matrix_elements, include/ed/ops/program.h:130-161, src/gpu/rep_matrix_elements.cu:175-340):
// out[p][o] = <bra_p| O_o |ket_p> = sum_r sum_c conj(bra_p[r]) M_o[r][c] ket_p[c], M_o the operator's
// matrix from the ket sector to the bra sector. One thread per bra representative walks the merged
// program once: per flip group one target and one lookup in the ket sector, then every observable's
// terms of the matching value subgroup summed to its h and its entry (the CSR build's entry) weighted
// into the pairs. The sums go to per-workgroup LDS accumulators (f64 atomics: their order within a
// workgroup is not fixed, so results agree to rounding, not bit for bit), then the workgroups' parts
// are added in workgroup order.
#include "dev_common.h"
#include "launch.h"

namespace qsa::dev {

namespace {

constexpr int kObsThreads = 256;

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

template <bool SLC>
__global__ __launch_bounds__(kObsThreads) void k_obs_sweep(SectorView R, SectorView C, int same, ProgramView P,
                                                           const u32* term_obs, int o0, int nob, int n_pairs,
                                                           const Zd* const* bras, const Zd* const* kets,
                                                           double* partial) {
    extern __shared__ double acc[];   // [n_pairs][nob] x (re, im)
    const int cells = n_pairs * nob;
    for (int i = threadIdx.x; i < 2 * cells; i += blockDim.x) acc[i] = 0.0;
    __syncthreads();
    for (i64 r = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; r < R.dim;
         r += static_cast<i64>(gridDim.x) * blockDim.x) {
        const u64 s = R.reps[r];
        const double w = R.inv_norms[r];
        for (u32 g = 0; g < P.n_groups; ++g) {
            const i64 vs = subgroup_of(P, s, g);
            if (vs < 0) continue;
            const u32 k0 = P.vsub_tbegin[vs], k1 = P.vsub_tbegin[vs + 1];
            if (term_obs[k0] >= static_cast<u32>(o0 + nob) || term_obs[k1 - 1] < static_cast<u32>(o0)) continue;
            const u64 t = s ^ P.group_flip[g];
            const bool diag = same && t == s;
            i64 c = r;
            double pr = 0.0, pi = 0.0;
            if (!diag) {
                c = index_and_projection<SLC>(C, t, pr, pi);
                if (c < 0) continue;
            }
            u32 k = k0;
            while (k < k1) {
                const u32 o = term_obs[k];
                double hr = 0.0, hi = 0.0;
                for (; k < k1 && term_obs[k] == o; ++k) {
                    if (popc(s & P.term_sign[k]) & 1) {
                        hr += -P.coeff_re[k];
                        hi += -P.coeff_im[k];
                    } else {
                        hr += P.coeff_re[k];
                        hi += P.coeff_im[k];
                    }
                }
                if (o < static_cast<u32>(o0) || o >= static_cast<u32>(o0 + nob)) continue;
                double vr, vi;   // the matrix entry M_o[r][c], as the CSR build writes it
                if (diag) {
                    vr = hr;
                    vi = -hi;
                } else {
                    const double xr = hr * pr - hi * pi;
                    const double xi = hr * pi + hi * pr;
                    vr = w * xr;
                    vi = w * (-xi);
                }
                for (int p = 0; p < n_pairs; ++p) {
                    const Zd b = bras[p][r], kv = kets[p][c];
                    // conj(b) v kv
                    const double mr = vr * kv.re - vi * kv.im, mi = vr * kv.im + vi * kv.re;
                    const double er = b.re * mr + b.im * mi, ei = b.re * mi - b.im * mr;
                    double* cell = acc + 2 * (p * nob + static_cast<int>(o) - o0);
                    atomicAdd(cell, er);
                    atomicAdd(cell + 1, ei);
                }
            }
        }
    }
    __syncthreads();
    double* out = partial + static_cast<u64>(blockIdx.x) * 2 * cells;
    for (int i = threadIdx.x; i < 2 * cells; i += blockDim.x) out[i] = acc[i];
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
    return static_cast<int>(b < 1 ? 1 : (b > 1024 ? 1024 : b));
}

void obs_sweep(const SectorView& R, const SectorView& C, bool same, const ProgramView& P, const u32* term_obs, int o0,
               int nob, int n_obs, int n_pairs, const Zd* const* bras, const Zd* const* kets, double* partial,
               double* out, hipStream_t st) {
    const int blocks = obs_sweep_blocks(R.dim);
    const std::size_t lds = static_cast<std::size_t>(n_pairs) * nob * 2 * sizeof(double);
    const bool walk = R.dim > 0 && P.n_groups > 0;
    if (walk) {
        if (C.slc_L > 0)
            hipLaunchKernelGGL(k_obs_sweep<true>, dim3(blocks), dim3(kObsThreads), lds, st, R, C, same ? 1 : 0, P,
                               term_obs, o0, nob, n_pairs, bras, kets, partial);
        else
            hipLaunchKernelGGL(k_obs_sweep<false>, dim3(blocks), dim3(kObsThreads), lds, st, R, C, same ? 1 : 0, P,
                               term_obs, o0, nob, n_pairs, bras, kets, partial);
    }
    const int cells = n_pairs * nob;
    hipLaunchKernelGGL(k_obs_reduce, dim3((cells + 255) / 256), dim3(256), 0, st, partial, walk ? blocks : 0, n_pairs,
                       nob, n_obs, o0, out);
}

}  // namespace qsa::dev
