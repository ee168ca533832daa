// radix_sort.hip -- stable LSD radix sort of (uint32 key, uint32 value) pairs
// for gfx950.  Replaces thrust::sort_by_key(hash, ids)
// (SPH2S_FINAL_SINGLE_SCALE/Solver.cu:1110-1115).
//
// One pass per 8-bit digit (fewer bits in the last pass), reduce-then-scan:
//   upsweep   per-tile digit histogram in LDS -> hist[digit][tile]
//   scan      exclusive scan of hist (digit-major, so tiles of one digit are
//             contiguous and earlier tiles come first: this is what makes the
//             sort stable across tiles)
//   downsweep re-reads the tile, ranks each key among equal digits with
//             wave64 ballots (lanes in key order), scatters to
//             hist[digit][tile] + rank.
// A tile is 4096 keys: 4 waves x 16 rounds x 64 lanes; wave w owns a
// contiguous run of 1024 keys so the in-tile order is (wave, round, lane).
#include "internal.hpp"

namespace sph {
namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kItems = 16;
constexpr int kWaveKeys = 64 * kItems;
constexpr int kTile = kBlock * kItems;
constexpr int kMaxRadix = 256;
constexpr int kScanThreads = 1024;

// Lanes (among valid ones) whose digit equals this lane's, via one ballot per
// digit bit.
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool valid, int bits) {
    uint64_t m = __ballot(valid);
    for (int b = 0; b < bits; ++b) {
        const bool set = (d >> b) & 1u;
        const uint64_t bm = __ballot(set);
        m &= set ? bm : ~bm;
    }
    return m;
}

__device__ __forceinline__ uint64_t lanemask_lt(int lane) {
    return (lane == 0) ? 0ull : (~0ull >> (64 - lane));
}

__global__ __launch_bounds__(kBlock) void k_upsweep(const uint32_t* __restrict__ keys,
                                                    uint32_t n, int shift, int bits,
                                                    uint32_t* __restrict__ hist,
                                                    uint32_t num_tiles) {
    __shared__ uint32_t cnt[kMaxRadix];
    const int radix = 1 << bits;
    for (int i = threadIdx.x; i < radix; i += kBlock) cnt[i] = 0;
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t tile = blockIdx.x;
    const uint32_t base = tile * kTile + wave * kWaveKeys;
    const uint32_t mask = static_cast<uint32_t>(radix - 1);
    for (int r = 0; r < kItems; ++r) {
        const uint32_t idx = base + r * 64 + lane;
        const bool valid = idx < n;
        const uint32_t d = valid ? (keys[idx] >> shift) & mask : 0u;
        const uint64_t peers = match_digit(d, valid, bits);
        if (valid && lane == __builtin_ctzll(peers))
            atomicAdd(&cnt[d], static_cast<uint32_t>(__popcll(peers)));
    }
    __syncthreads();
    for (int i = threadIdx.x; i < radix; i += kBlock) hist[i * num_tiles + tile] = cnt[i];
}

// Exclusive scan of len entries by one workgroup: each thread owns a
// contiguous chunk, chunk sums are scanned in LDS.
__global__ __launch_bounds__(kScanThreads) void k_scan(uint32_t* __restrict__ data,
                                                       uint32_t len) {
    __shared__ uint32_t sums[kScanThreads];
    const uint32_t t = threadIdx.x;
    const uint32_t chunk = (len + kScanThreads - 1) / kScanThreads;
    const uint32_t b = min(t * chunk, len), e = min(b + chunk, len);
    uint32_t s = 0;
    for (uint32_t i = b; i < e; ++i) s += data[i];
    sums[t] = s;
    __syncthreads();
    for (uint32_t off = 1; off < kScanThreads; off <<= 1) {
        const uint32_t v = (t >= off) ? sums[t - off] : 0u;
        __syncthreads();
        sums[t] += v;
        __syncthreads();
    }
    uint32_t run = (t > 0) ? sums[t - 1] : 0u;
    for (uint32_t i = b; i < e; ++i) {
        const uint32_t v = data[i];
        data[i] = run;
        run += v;
    }
}

__global__ __launch_bounds__(kBlock) void k_downsweep(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out, uint32_t n,
    int shift, int bits, const uint32_t* __restrict__ hist, uint32_t num_tiles) {
    __shared__ uint32_t wcnt[kWaves][kMaxRadix];
    __shared__ uint32_t gbase[kMaxRadix];
    const int radix = 1 << bits;
    const uint32_t mask = static_cast<uint32_t>(radix - 1);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t tile = blockIdx.x;
    for (int i = threadIdx.x; i < kWaves * kMaxRadix; i += kBlock) (&wcnt[0][0])[i] = 0u;
    for (int i = threadIdx.x; i < radix; i += kBlock) gbase[i] = hist[i * num_tiles + tile];
    __syncthreads();

    const uint32_t base = tile * kTile + wave * kWaveKeys;
    const uint64_t lt = lanemask_lt(lane);
    uint32_t key[kItems], val[kItems], rank[kItems];
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        const uint32_t idx = base + r * 64 + lane;
        const bool valid = idx < n;
        key[r] = valid ? keys_in[idx] : 0u;
        val[r] = valid ? (vals_in ? vals_in[idx] : idx) : 0u;
    }
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        const uint32_t idx = base + r * 64 + lane;
        const bool valid = idx < n;
        const uint32_t d = (key[r] >> shift) & mask;
        const uint64_t peers = match_digit(d, valid, bits);
        const uint32_t before = wcnt[wave][d];
        rank[r] = before + static_cast<uint32_t>(__popcll(peers & lt));
        // The read above is one wave instruction issued before this write, and
        // the written value depends on it, so every lane sees the pre-round count.
        if (valid && lane == __builtin_ctzll(peers))
            wcnt[wave][d] = before + static_cast<uint32_t>(__popcll(peers));
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // Per digit: exclusive prefix over waves (in-tile order is wave-major).
    for (int d = threadIdx.x; d < radix; d += kBlock) {
        uint32_t run = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t c = wcnt[w][d];
            wcnt[w][d] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        const uint32_t idx = base + r * 64 + lane;
        if (idx < n) {
            const uint32_t d = (key[r] >> shift) & mask;
            const uint32_t dst = gbase[d] + wcnt[wave][d] + rank[r];
            if (dst < n) {  // always true for a consistent histogram; guards the scatter
                keys_out[dst] = key[r];
                vals_out[dst] = val[r];
            }
        }
    }
}

inline uint32_t tiles_for(uint32_t n) { return (n + kTile - 1) / kTile; }

}  // namespace

size_t radix_sort_temp_bytes(uint32_t n) {
    const size_t tiles = tiles_for(n > 0 ? n : 1);
    const size_t hist = static_cast<size_t>(kMaxRadix) * tiles * sizeof(uint32_t);
    const size_t hist_aligned = (hist + 255) & ~static_cast<size_t>(255);
    const size_t buf = (static_cast<size_t>(n) * sizeof(uint32_t) + 255) & ~static_cast<size_t>(255);
    return hist_aligned + 2 * buf;
}

int radix_sort_pairs(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
                     uint32_t* vals_out, uint32_t n, uint32_t key_bits, void* temp,
                     size_t temp_bytes, hipStream_t stream) {
    if (key_bits < 1 || key_bits > 32)
        return set_error(SPH_ERR_INVALID_ARGUMENT, "radix sort: key_bits must be in [1, 32]");
    if (n == 0) return SPH_OK;
    if (!keys_in || !keys_out || !vals_out)
        return set_error(SPH_ERR_INVALID_ARGUMENT, "radix sort: null buffer");
    if (keys_in == keys_out || vals_in == vals_out)
        return set_error(SPH_ERR_INVALID_ARGUMENT, "radix sort: outputs alias inputs");
    if (!temp || temp_bytes < radix_sort_temp_bytes(n))
        return set_error(SPH_ERR_INVALID_ARGUMENT, "radix sort: temp buffer too small");
    const uint32_t tiles = tiles_for(n);
    const size_t hist_bytes = static_cast<size_t>(kMaxRadix) * tiles * sizeof(uint32_t);
    const size_t hist_aligned = (hist_bytes + 255) & ~static_cast<size_t>(255);
    const size_t buf = (static_cast<size_t>(n) * sizeof(uint32_t) + 255) & ~static_cast<size_t>(255);
    auto* hist = static_cast<uint32_t*>(temp);
    auto* ktmp = reinterpret_cast<uint32_t*>(static_cast<char*>(temp) + hist_aligned);
    auto* vtmp = reinterpret_cast<uint32_t*>(static_cast<char*>(temp) + hist_aligned + buf);

    const int passes = static_cast<int>((key_bits + 7) / 8);
    const uint32_t* kin = keys_in;
    const uint32_t* vin = vals_in;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        const int bits = static_cast<int>(key_bits) - shift < 8 ? static_cast<int>(key_bits) - shift : 8;
        // The last pass must land in the caller's output buffers.
        const bool to_out = ((passes - 1 - p) % 2) == 0;
        uint32_t* kout = to_out ? keys_out : ktmp;
        uint32_t* vout = to_out ? vals_out : vtmp;
        hipLaunchKernelGGL(k_upsweep, dim3(tiles), dim3(kBlock), 0, stream, kin, n, shift, bits,
                           hist, tiles);
        hipLaunchKernelGGL(k_scan, dim3(1), dim3(kScanThreads), 0, stream, hist,
                           static_cast<uint32_t>(1u << bits) * tiles);
        hipLaunchKernelGGL(k_downsweep, dim3(tiles), dim3(kBlock), 0, stream, kin, vin, kout,
                           vout, n, shift, bits, hist, tiles);
        kin = kout;
        vin = vout;
    }
    SPH_HIP(hipGetLastError());
    return SPH_OK;
}

}  // namespace sph
