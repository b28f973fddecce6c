// FTLM start vectors drawn on the device.
//
// The reference draws every sample vector on the host: std::mt19937 seeded with the sample's
// 32-bit seed, then entry i = (n_{2i}, n_{2i+1}) from one std::normal_distribution<double>(0, 1)
// (include/ed/thermal/ftlm.h, sample_seed.h:37-89). With libstdc++ (GCC 11, bits/random.tcc) that
// distribution is Marsaglia's polar method on 53-bit canonical uniforms, u = (w0 + 2^32 w1) / 2^64
// from two consecutive 32-bit outputs w0, w1 (rounded once, clamped below 1): a candidate pair
// x = 2 u_a - 1, y = 2 u_b - 1 takes four outputs, is kept when 0 < r2 = x^2 + y^2 <= 1, and gives
// the two normals y m, x m with m = sqrt(-2 ln r2 / r2), in that order. A kept candidate is thus one
// complex entry (y m, x m), and the stream of entries is the stream of kept candidates.
//
// One workgroup per vector. Its 624-word Mersenne Twister state lives in LDS (two copies); one
// regeneration of the state is three dependent ranges of the recurrence (words [0, 227) from the old
// state, [227, 454) and [454, 624) from new words 227 positions back), each range in parallel. The
// 624 tempered outputs are exactly 156 candidates; a ballot per wave and the popcounts of the waves
// before give each kept candidate its entry index. Everything but ln and sqrt is exact integer or
// correctly rounded IEEE arithmetic in the library's order, so the entries equal the host's up to
// the last bit of the logarithm.
#include "dev_common.h"
#include "launch.h"

namespace qsa::dev {

namespace {

constexpr int kMtN = 624;
constexpr int kMtM = 397;
constexpr int kSampleThreads = 256;
constexpr int kCandidates = kMtN / 4;   // 156: four outputs per candidate pair

__device__ __forceinline__ u32 mt_twist(u32 cur, u32 next) {
    const u32 y = (cur & 0x80000000u) | (next & 0x7fffffffu);
    return (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
}

__device__ __forceinline__ u32 mt_temper(u32 y) {
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

// generate_canonical<double, 53> on two 32-bit outputs.
__device__ __forceinline__ double canonical(u32 lo, u32 hi) {
    double sum = static_cast<double>(lo);
    sum = sum + static_cast<double>(hi) * 4294967296.0;
    const double r = sum / 18446744073709551616.0;
    return r >= 1.0 ? 0x1.fffffffffffffp-1 : r;
}

// Vector p of a batch: block b = p / R, slot s = p % R; entries go to V[(row_begin[b] + i) * R + s]
// for i < dims[b].
__global__ __launch_bounds__(kSampleThreads) void k_sample_gauss(Zd* V, int R, const u64* row_begin, const u64* dims,
                                                                 const u32* seeds) {
    __shared__ u32 state[2][kMtN];
    __shared__ u64 kept[kSampleThreads / 64];
    const u32 p = blockIdx.x;
    const u32 b = p / static_cast<u32>(R);
    const int s = static_cast<int>(p % static_cast<u32>(R));
    const u64 n = dims[b];
    const u64 row0 = row_begin[b];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) {
        u32 x = seeds[p];
        state[0][0] = x;
        for (u32 i = 1; i < kMtN; ++i) {
            x = 1812433253u * (x ^ (x >> 30)) + i;
            state[0][i] = x;
        }
    }
    __syncthreads();
    int cur = 0;
    u64 done = 0;
    while (done < n) {
        const u32* o = state[cur];
        u32* w = state[cur ^ 1];
        for (int i = tid; i < kMtN - kMtM; i += kSampleThreads) w[i] = o[i + kMtM] ^ mt_twist(o[i], o[i + 1]);
        __syncthreads();
        for (int i = kMtN - kMtM + tid; i < 2 * (kMtN - kMtM); i += kSampleThreads)
            w[i] = w[i - (kMtN - kMtM)] ^ mt_twist(o[i], o[i + 1]);
        __syncthreads();
        for (int i = 2 * (kMtN - kMtM) + tid; i < kMtN; i += kSampleThreads)
            w[i] = w[i - (kMtN - kMtM)] ^ mt_twist(o[i], i + 1 < kMtN ? o[i + 1] : w[0]);
        __syncthreads();
        cur ^= 1;
        bool keep = false;
        double re = 0.0, im = 0.0;
        if (tid < kCandidates) {
            const double ua = canonical(mt_temper(w[4 * tid]), mt_temper(w[4 * tid + 1]));
            const double ub = canonical(mt_temper(w[4 * tid + 2]), mt_temper(w[4 * tid + 3]));
            const double x = 2.0 * ua - 1.0;
            const double y = 2.0 * ub - 1.0;
            const double r2 = x * x + y * y;
            keep = !(r2 > 1.0 || r2 == 0.0);
            if (keep) {
                const double m = sqrt(-2.0 * log(r2) / r2);
                im = y * m;   // the pair's first normal is the entry's imaginary part (the reference's
                re = x * m;   // compiler evaluates Complex(nd(gen), nd(gen))'s second argument first)
            }
        }
        const u64 mask = __ballot(keep);
        if (lane == 0) kept[wave] = mask;
        __syncthreads();
        u64 before = static_cast<u64>(__popcll(mask & ((u64{1} << lane) - 1)));
        u64 total = 0;
        for (int wv = 0; wv < kSampleThreads / 64; ++wv) {
            const u64 c = static_cast<u64>(__popcll(kept[wv]));
            if (wv < wave) before += c;
            total += c;
        }
        if (keep && done + before < n) V[(row0 + done + before) * static_cast<u64>(R) + s] = Zd{re, im};
        done += total;
        __syncthreads();
    }
}

}  // namespace

void sample_gauss(Zd* V, int R, const u64* row_begin, const u64* dims, const u32* seeds, u32 vectors, hipStream_t st) {
    if (!vectors) return;
    hipLaunchKernelGGL(k_sample_gauss, dim3(vectors), dim3(kSampleThreads), 0, st, V, R, row_begin, dims, seeds);
}

}  // namespace qsa::dev
