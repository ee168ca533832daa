// radix_sort.hip -- stable LSD radix sort of (uint32 key, uint32 value) pairs
// for gfx950.  Replaces thrust::sort_by_key(hash, ids)
// (SPH2S_FINAL_SINGLE_SCALE/Solver.cu:1110-1115).
//
// Single-sweep scheme, 10-bit digits (fewer bits in the last pass; 19-bit
// cell hashes take two passes), one thread per digit in the pass kernel:
//   k_prepare   zeroes the digit histograms, tile tickets and look-back
//               status words (and, for the solver, resets the cell ranges)
//   k_hist      one read of the keys -> LDS digit histograms of every pass
//               (counted by runs of equal digits: the solver's keys arrive
//               nearly sorted), added into the global per-pass histograms
//   k_pass      one kernel per digit: a block takes a tile ticket (so tiles
//               start in order), ranks its keys among equal digits with wave64
//               ballots, publishes its per-digit counts, looks back over the
//               status words of earlier tiles for its exclusive prefix
//               (decoupled look-back, kLookBack tiles per round trip), and
//               scatters.
// A tile is kPassBlock*kItems keys (1024 threads x 4 keys: 16 waves share a
// 4096-key tile, so each wave ranks 4 rounds instead of 16 and the CU has 16
// waves to hide latency with): wave w owns a contiguous run of 64*kItems
// keys, so the in-tile order is (wave, round, lane) and the sort is stable.
// Every look-back wait is bounded: a wait that exceeds kSpinLimit polls sets
// an error bit and gives up instead of hanging.
#include <algorithm>
#include <mutex>

#include "neighbors.hpp"

namespace sph {
namespace {

constexpr int kBlock = 256;
// Digit passes: kPassBlock threads rank a tile of kPassBlock * kItems keys.
#ifndef SPH_SORT_BLOCK
#define SPH_SORT_BLOCK 1024
#endif
#ifndef SPH_SORT_ITEMS
#define SPH_SORT_ITEMS 4
#endif
constexpr int kPassBlock = SPH_SORT_BLOCK;
constexpr int kWaves = kPassBlock / 64;
static_assert(kPassBlock >= 256 && kPassBlock <= 1024 && kPassBlock % 64 == 0,
              "digit-pass block shape");
constexpr int kItems = SPH_SORT_ITEMS;
constexpr int kWaveKeys = 64 * kItems;
constexpr int kTile = kPassBlock * kItems;
// The histogram kernel reads its keys in smaller tiles (more blocks in
// flight): kHistItems keys per thread, set by each key source below.
#ifndef SPH_SORT_DIGIT_BITS
#define SPH_SORT_DIGIT_BITS 10
#endif
constexpr int kDigitBits = SPH_SORT_DIGIT_BITS;
constexpr int kRadix = 1 << kDigitBits;
constexpr int kMaxPasses = (32 + kDigitBits - 1) / kDigitBits;
constexpr uint32_t kFlagAggregate = 1u << 30;
constexpr uint32_t kFlagPrefix = 2u << 30;
constexpr uint32_t kCountMask = (1u << 30) - 1u;
constexpr uint32_t kSpinLimit = 1u << 22;
#ifndef SPH_SORT_LOOKBACK
#define SPH_SORT_LOOKBACK 4
#endif
constexpr int kLookBack = SPH_SORT_LOOKBACK;  // predecessor status words per look-back round trip

// Header of the temporary buffer, in 32-bit words.
constexpr uint32_t kHistWords = kMaxPasses * kRadix;        // global histograms
constexpr uint32_t kCounterWords = 64;                      // tickets + error word
constexpr uint32_t kHeaderWords = kHistWords + kCounterWords;
constexpr uint32_t kErrorWord = kHeaderWords - 1;           // never zeroed by k_prepare
constexpr uint32_t kNWord = kHeaderWords - 2;               // live key count, set by k_prepare

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

__device__ __forceinline__ uint32_t load_status(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void store_status(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Also publishes the live key count (from n_src, or n_host when n_src is
// null) in the header word kNWord, which every later sort kernel reads.
__global__ __launch_bounds__(kBlock) void k_prepare(uint32_t* __restrict__ zero, uint32_t zero_words,
                                                    uint32_t* __restrict__ fill_a,
                                                    uint32_t* __restrict__ fill_b,
                                                    uint32_t fill_words,
                                                    const uint32_t* __restrict__ n_src,
                                                    uint32_t n_host) {
    const uint32_t stride = gridDim.x * kBlock;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < zero_words; i += stride) {
        if (i == kNWord)
            zero[i] = n_src ? *n_src : n_host;
        else if (i != kErrorWord)
            zero[i] = 0u;
    }
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < fill_words; i += stride) {
        fill_a[i] = kEmptyCell;
        fill_b[i] = kEmptyCell;
    }
}

// Keys read from an array.
struct ArrayKeys {
    static constexpr int kHistItems = 8;
    const uint32_t* keys;
    __device__ uint32_t operator()(uint32_t i) const { return keys[i]; }
};

// Keys computed as cell hashes of float4 positions (computeHashs,
// SPH2S_FINAL_SINGLE_SCALE/Solver.cu:401-424), written out for the first pass.
// kHistItems: 4 keys per thread (1,024-key blocks, 977 blocks for 1M keys)
// took the hash + histogram pass of the 1M-particle step from 19.4 to 16.8 us
// against 8; 2 was as fast, 1, 16 and 32 slower.
struct HashKeys {
    static constexpr int kHistItems = 4;
    const float4* pos;
    uint32_t* keys_out;
    float ox, oy, oz, spacing;
    int dx, dy, dz;
    __device__ uint32_t operator()(uint32_t i) const {
        const float4 x = pos[i];
        const uint32_t k = cell_hash(x.x, x.y, x.z, ox, oy, oz, spacing, dx, dy, dz);
        keys_out[i] = k;
        return k;
    }
};

// Batched scenes: the scene id is in vel.w; the key is the scene's first global
// cell plus the reference hash of the position in the scene's own grid, so a
// stable sort keeps every scene contiguous and in its own hash order.
// (kHistItems 8: with 4 the 4.1M-key batch of 64 scenes ran its step 1%
// slower.)
struct BatchHashKeys {
    static constexpr int kHistItems = 8;
    const float4* pos;
    const float4* vel;
    const SceneParams* scenes;
    uint32_t* keys_out;
    __device__ uint32_t operator()(uint32_t i) const {
        const float4 x = pos[i];
        const GridView& g = scenes[__float_as_uint(vel[i].w)].g;
        const uint32_t k =
            g.cell_offset + cell_hash(x.x, x.y, x.z, g.ox, g.oy, g.oz, g.spacing, g.dx, g.dy, g.dz);
        keys_out[i] = k;
        return k;
    }
};

// The keys arrive nearly sorted (the solver's particles are stored in the
// previous step's cell order), so each digit is counted by runs: a lane whose
// digit differs from the previous lane's starts a run and adds its length.
// The keys of all rounds are computed (and their positions loaded) before any
// is counted.
template <class Src>
__global__ __launch_bounds__(kBlock) void k_hist(Src src, const uint32_t* __restrict__ n_dev,
                                                 int passes, uint32_t key_bits,
                                                 uint32_t* __restrict__ ghist) {
    constexpr int kHistItems = Src::kHistItems;
    constexpr uint32_t kHistTile = kBlock * kHistItems;
    const uint32_t n = *n_dev;
    if (blockIdx.x * kHistTile >= n) return;  // block-uniform: the grid covers the capacity
    __shared__ uint32_t cnt[kMaxPasses][kRadix];
    for (int i = threadIdx.x; i < kMaxPasses * kRadix; i += kBlock) (&cnt[0][0])[i] = 0u;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t base = blockIdx.x * kHistTile + wave * (64 * kHistItems);
    uint32_t key[kHistItems];
#pragma unroll
    for (int r = 0; r < kHistItems; ++r) {
        const uint32_t idx = base + r * 64 + lane;
        key[r] = (idx < n) ? src(idx) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kHistItems; ++r) {
        const uint32_t first = base + r * 64;
        // Valid lanes form a prefix of the wave.
        const uint32_t nvalid = (first < n) ? min(64u, n - first) : 0u;
        const bool valid = static_cast<uint32_t>(lane) < nvalid;
        const uint32_t prev = static_cast<uint32_t>(__shfl_up(static_cast<int>(key[r]), 1));
        for (int p = 0; p < passes; ++p) {
            const int shift = kDigitBits * p;
            const int bits = min(kDigitBits, static_cast<int>(key_bits) - shift);
            const uint32_t m = (1u << bits) - 1u;
            const uint32_t d = (key[r] >> shift) & m;
            const bool start = valid && (lane == 0 || ((prev >> shift) & m) != d);
            const uint64_t starts = __ballot(start);
            if (start) {
                const uint64_t later = (lane == 63) ? 0ull : (starts >> (lane + 1));
                const uint32_t end =
                    later ? static_cast<uint32_t>(lane + 1 + __builtin_ctzll(later)) : nvalid;
                atomicAdd(&cnt[p][d], end - static_cast<uint32_t>(lane));
            }
        }
    }
    __syncthreads();
    for (int p = 0; p < passes; ++p) {
        const int bits = min(kDigitBits, static_cast<int>(key_bits) - kDigitBits * p);
        for (int d = threadIdx.x; d < (1 << bits); d += kBlock)
            if (cnt[p][d]) atomicAdd(&ghist[p * kRadix + d], cnt[p][d]);
    }
}

__global__ __launch_bounds__(kPassBlock) void k_pass(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out,
    const uint32_t* __restrict__ n_dev, int shift, int bits, const uint32_t* __restrict__ ghist,
    uint32_t* __restrict__ status, uint32_t* __restrict__ ticket,
    uint32_t* __restrict__ error_word) {
    __shared__ uint32_t s_tile;
    // Per-wave digit counts, then their exclusive prefix over the waves: at
    // most one tile, kTile <= 65535, so 16 bits suffice.
    static_assert(kRadix <= kPassBlock, "a digit pass needs one thread per digit");
    static_assert(kTile <= 65535, "16-bit digit counts");
    __shared__ uint16_t wcnt[kWaves][kRadix];
    __shared__ uint32_t dbase[kRadix];
    const int radix = 1 << bits;
    const uint32_t mask = static_cast<uint32_t>(radix - 1);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int t = threadIdx.x;

    if (t == 0) s_tile = atomicAdd(ticket, 1u);
    for (int i = t; i < kWaves * kRadix; i += kPassBlock) (&wcnt[0][0])[i] = 0;
    // Exclusive scan of this pass's global digit histogram (threads t < 256).
    if (t < kRadix) dbase[t] = (t < radix) ? ghist[t] : 0u;
    __syncthreads();
    for (int off = 1; off < kRadix; off <<= 1) {
        const uint32_t v = (t < kRadix && t >= off) ? dbase[t - off] : 0u;
        __syncthreads();
        if (t < kRadix) dbase[t] += v;
        __syncthreads();
    }
    const uint32_t digit_excl = (t < radix) ? dbase[t] - ghist[t] : 0u;
    const uint32_t tile = s_tile;
    // The grid covers the capacity; tickets past the live tiles have no keys.
    // Tickets are taken in dispatch order, so live tiles never wait on them.
    const uint32_t n = *n_dev;
    if (tile * kTile >= n) return;

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
        // One wave instruction reads the count before the leader's write, and
        // the written value depends on the read, so all lanes see the old count.
        if (valid && lane == __builtin_ctzll(peers))
            wcnt[wave][d] = static_cast<uint16_t>(before + static_cast<uint32_t>(__popcll(peers)));
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    if (t < radix) {
        uint32_t run = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t c = wcnt[w][t];
            wcnt[w][t] = static_cast<uint16_t>(run);
            run += c;
        }
        // Decoupled look-back over earlier tiles for digit t, kLookBack
        // predecessors per round trip: their words are loaded together and
        // taken nearest first -- aggregates are added until an inclusive
        // prefix ends the walk, and a word not yet published stops the window
        // (it is polled again).  Tile 0 always publishes a prefix.
        uint32_t* my = status + static_cast<size_t>(tile) * kRadix + t;
        uint32_t excl = 0;
        if (tile == 0) {
            store_status(my, kFlagPrefix | run);
        } else {
            store_status(my, kFlagAggregate | run);
            int32_t pred = static_cast<int32_t>(tile) - 1;
            uint32_t spins = 0;
            bool done = false;
            while (!done) {
                uint32_t v[kLookBack];
#pragma unroll
                for (int q = 0; q < kLookBack; ++q)
                    v[q] = (pred - q >= 0)
                               ? load_status(status + static_cast<size_t>(pred - q) * kRadix + t)
                               : 0u;
                int taken = 0;
#pragma unroll
                for (int q = 0; q < kLookBack; ++q) {
                    if (done || taken < q || pred - q < 0) continue;
                    const uint32_t flag = v[q] & ~kCountMask;
                    if (flag == 0u) continue;  // not published: the window stops here
                    excl += v[q] & kCountMask;
                    taken = q + 1;
                    done = flag == kFlagPrefix;
                }
                if (done) break;
                if (taken == 0) {
                    if (++spins > kSpinLimit) {
                        atomicOr(error_word, kDevErrSortWait);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                pred -= taken;
                if (pred < 0) break;  // unreachable: tile 0 publishes a prefix
            }
            store_status(my, kFlagPrefix | ((excl + run) & kCountMask));
        }
        dbase[t] = digit_excl + excl;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        const uint32_t idx = base + r * 64 + lane;
        if (idx < n) {
            const uint32_t d = (key[r] >> shift) & mask;
            const uint32_t dst = dbase[d] + wcnt[wave][d] + rank[r];
            if (dst < n) {  // always true for consistent counts; guards the scatter
                keys_out[dst] = key[r];
                vals_out[dst] = val[r];
            }
        }
    }
}

constexpr uint32_t tiles_for(uint32_t n) { return (n + kTile - 1) / kTile; }
constexpr size_t align_up(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

// Temporary bytes to sort n pairs: header, look-back status words, and the
// ping-pong keys and values of the intermediate passes.
constexpr size_t temp_bytes_for(uint32_t n) {
    return align_up(kHeaderWords * sizeof(uint32_t)) +
           align_up(static_cast<size_t>(kMaxPasses) * tiles_for(n > 0 ? n : 1) * kRadix *
                    sizeof(uint32_t)) +
           2 * align_up(static_cast<size_t>(n) * sizeof(uint32_t));
}

struct TempLayout {
    uint32_t* header;   // ghist[kMaxPasses][256], tickets, error word
    uint32_t* status;   // [passes][tiles][256]
    uint32_t* ktmp;
    uint32_t* vtmp;
    size_t status_words;
};

TempLayout layout(void* temp, uint32_t n) {
    TempLayout L;
    const size_t tiles = tiles_for(n > 0 ? n : 1);
    char* p = static_cast<char*>(temp);
    L.header = reinterpret_cast<uint32_t*>(p);
    p += align_up(kHeaderWords * sizeof(uint32_t));
    L.status = reinterpret_cast<uint32_t*>(p);
    L.status_words = static_cast<size_t>(kMaxPasses) * tiles * kRadix;
    p += align_up(L.status_words * sizeof(uint32_t));
    L.ktmp = reinterpret_cast<uint32_t*>(p);
    p += align_up(static_cast<size_t>(n) * sizeof(uint32_t));
    L.vtmp = reinterpret_cast<uint32_t*>(p);
    return L;
}

int check_sort_args(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
                    uint32_t* vals_out, uint32_t n, uint32_t key_bits, void* temp,
                    size_t temp_bytes) {
    if (key_bits < 1 || key_bits > 32)
        return set_error(SPH_ERR_INVALID_ARGUMENT, "radix sort: key_bits must be in [1, 32]");
    if (n >= (1u << 30)) return set_error(SPH_ERR_INVALID_ARGUMENT, "radix sort: n >= 2^30");
    if (n == 0) return SPH_OK;
    if (!keys_out || !vals_out)
        return set_error(SPH_ERR_INVALID_ARGUMENT, "radix sort: null buffer");
    if (keys_in == keys_out || (vals_in && vals_in == vals_out))
        return set_error(SPH_ERR_INVALID_ARGUMENT, "radix sort: outputs alias inputs");
    if (!temp || temp_bytes < radix_sort_temp_bytes(n))
        return set_error(SPH_ERR_INVALID_ARGUMENT, "radix sort: temp buffer too small");
    return SPH_OK;
}

// k_prepare + k_hist: zero the sort state (and reset the cell ranges), then
// build the digit histograms of every pass in one read of the keys.  cap is
// the launch capacity; the live count comes from n_src (device) or n_host.
template <class Src>
int prepare_hist(Src src, uint32_t cap, const uint32_t* n_src, uint32_t n_host, uint32_t key_bits,
                 void* temp, uint32_t* fill_a, uint32_t* fill_b, uint32_t fill_words,
                 hipStream_t stream) {
    const uint32_t tiles = tiles_for(cap);
    const TempLayout L = layout(temp, cap);
    const int passes = static_cast<int>((key_bits + kDigitBits - 1) / kDigitBits);
    const uint32_t zero_words =
        static_cast<uint32_t>(L.status - L.header) + static_cast<uint32_t>(passes) * tiles * kRadix;
    const uint32_t prep_words = zero_words > fill_words ? zero_words : fill_words;
    const uint32_t prep_blocks = std::min<uint32_t>((prep_words + kBlock - 1) / kBlock, 2048u);
    hipLaunchKernelGGL(k_prepare, dim3(prep_blocks), dim3(kBlock), 0, stream, L.header, zero_words,
                       fill_a, fill_b, fill_words, n_src, n_host);
    constexpr uint32_t kHistTile = kBlock * Src::kHistItems;
    const uint32_t hist_blocks = std::max(1u, (cap + kHistTile - 1) / kHistTile);
    hipLaunchKernelGGL(k_hist<Src>, dim3(hist_blocks), dim3(kBlock), 0, stream, src,
                       L.header + kNWord, passes, key_bits, L.header);
    SPH_HIP(hipGetLastError());
    return SPH_OK;
}

// One k_pass launch per digit; the last pass writes keys_out/vals_out.
int run_passes(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
               uint32_t* vals_out, uint32_t cap, uint32_t key_bits, void* temp,
               hipStream_t stream) {
    const uint32_t tiles = tiles_for(cap);
    const TempLayout L = layout(temp, cap);
    const int passes = static_cast<int>((key_bits + kDigitBits - 1) / kDigitBits);
    const uint32_t* kin = keys_in;
    const uint32_t* vin = vals_in;
    for (int p = 0; p < passes; ++p) {
        const int shift = kDigitBits * p;
        const int bits = std::min(kDigitBits, static_cast<int>(key_bits) - shift);
        const bool to_out = ((passes - 1 - p) % 2) == 0;
        uint32_t* kout = to_out ? keys_out : L.ktmp;
        uint32_t* vout = to_out ? vals_out : L.vtmp;
        hipLaunchKernelGGL(k_pass, dim3(tiles), dim3(kPassBlock), 0, stream, kin, vin, kout, vout,
                           L.header + kNWord, shift, bits, L.header + p * kRadix,
                           L.status + static_cast<size_t>(p) * tiles * kRadix,
                           L.header + kHistWords + p, L.header + kErrorWord);
        kin = kout;
        vin = vout;
    }
    SPH_HIP(hipGetLastError());
    return SPH_OK;
}

}  // namespace

size_t radix_sort_temp_bytes(uint32_t n) { return temp_bytes_for(n); }

uint32_t* radix_sort_error_word(void* temp) {
    return static_cast<uint32_t*>(temp) + kErrorWord;
}

int radix_sort_pairs(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
                     uint32_t* vals_out, uint32_t n, uint32_t key_bits, void* temp,
                     size_t temp_bytes, hipStream_t stream) {
    int rc = check_sort_args(keys_in, vals_in, keys_out, vals_out, n, key_bits, temp, temp_bytes);
    if (rc != SPH_OK || n == 0) return rc;
    if (!keys_in) return set_error(SPH_ERR_INVALID_ARGUMENT, "radix sort: null keys");
    rc = prepare_hist(ArrayKeys{keys_in}, n, nullptr, n, key_bits, temp, nullptr, nullptr, 0u,
                      stream);
    if (rc != SPH_OK) return rc;
    return run_passes(keys_in, vals_in, keys_out, vals_out, n, key_bits, temp, stream);
}

int hash_hist_cells(const StepParams& p, const uint32_t* n_dev, const float4* pos, uint32_t* keys,
                    uint32_t key_bits, void* temp, size_t temp_bytes, uint32_t* cell_start,
                    uint32_t* cell_end, hipStream_t stream) {
    if (!temp || temp_bytes < radix_sort_temp_bytes(p.n))
        return set_error(SPH_ERR_INVALID_ARGUMENT, "hash: sort temp buffer too small");
    HashKeys src{pos, keys, p.ox, p.oy, p.oz, p.spacing, p.dx, p.dy, p.dz};
    return prepare_hist(src, p.n, n_dev, p.n, key_bits, temp, cell_start, cell_end, p.num_cells,
                        stream);
}

int hash_hist_cells_batched(const StepParams& cap, const uint32_t* n_dev, const float4* pos,
                            const float4* vel, const SceneParams* scenes, uint32_t* keys,
                            uint32_t key_bits, void* temp, size_t temp_bytes,
                            uint32_t* cell_start, uint32_t* cell_end, hipStream_t stream) {
    if (!temp || temp_bytes < radix_sort_temp_bytes(cap.n))
        return set_error(SPH_ERR_INVALID_ARGUMENT, "hash: sort temp buffer too small");
    BatchHashKeys src{pos, vel, scenes, keys};
    return prepare_hist(src, cap.n, n_dev, cap.n, key_bits, temp, cell_start, cell_end,
                        cap.num_cells, stream);
}

int sort_cells(const StepParams& p, const uint32_t* keys, uint32_t* skeys, uint32_t* perm,
               uint32_t key_bits, void* temp, hipStream_t stream) {
    return run_passes(keys, nullptr, skeys, perm, p.n, key_bits, temp, stream);
}

// ---------------------------------------------------------------------------
// The sort without any allocation: the temporary buffer is a file-scope
// device array of fixed size, one copy per GPU for the life of the process.
// Host code never passes the array's name to a kernel (on the host it only
// stands in for the symbol): the first call on a device takes the array's
// device address with hipGetSymbolAddress, with that device current, and
// every later call on that device uses the kept address.
constexpr size_t kStaticSortBytes = temp_bytes_for(kStaticSortMaxKeys);  // 24.0 MiB
static __device__ __align__(256) unsigned char scratch[kStaticSortBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch[kMaxDevices] = {};

// The device address of this GPU's scratch array (taken once per device).
int scratch_address(void** out) {
    int dev = 0;
    SPH_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return set_error(SPH_ERR_UNSUPPORTED, "radix sort: device ordinal beyond 63");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return hip_error(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress",
                             __FILE__, __LINE__);
        g_scratch[dev] = ptr;
    }
    *out = g_scratch[dev];
    return SPH_OK;
}

}  // namespace

int radix_sort_pairs_static(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
                            uint32_t* vals_out, uint32_t n, uint32_t key_bits,
                            hipStream_t stream) {
    if (n > kStaticSortMaxKeys)
        return set_error(SPH_ERR_CAPACITY,
                         "radix sort without allocation: n exceeds its maximum of 2^21 pairs "
                         "(use sph_radix_sort_pairs with a caller-allocated buffer)");
    if (n == 0)
        return check_sort_args(keys_in, vals_in, keys_out, vals_out, n, key_bits, nullptr, 0);
    void* temp = nullptr;
    const int rc = scratch_address(&temp);
    if (rc != SPH_OK) return rc;
    return radix_sort_pairs(keys_in, vals_in, keys_out, vals_out, n, key_bits, temp,
                            kStaticSortBytes, stream);
}

}  // namespace sph
