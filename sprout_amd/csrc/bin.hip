// bin.hip — binning of the particles by cell on gfx950 (sort.comp, claim.comp), written for
// wave64 and LDS: no library sort.
//
// The reference counts particles per cell with global atomics, prefix-sums the counts with
// one thread per row and then one thread over the rows, and scatters with an atomic per
// particle (sort.comp:26-68), so the order inside a cell is whatever order the atomics ran
// in. Here the order is defined: stable by slot. The sort is an LSD radix sort of 32-bit
// (key, slot) pairs, 8 bits per pass:
//   1. histogram: each group counts the digits of its BIN_TILE keys in LDS; a wave's lanes
//      holding the same digit are found with eight ballots (their "peers") and the lowest of
//      them adds their number, one LDS atomic per distinct digit per wave;
//   2. scan: the digit-major histogram (digit * groups + group) is prefix-summed, chunk by
//      chunk in LDS (wave inclusive scans with DPP shuffles, then the waves' totals), then
//      over the chunks' totals by one group; the sum over earlier digits and earlier groups
//      is where a group's keys of one digit start;
//   3. scatter: each group walks its keys in slot order, 256 at a time; a key's place is its
//      digit's start + the keys of that digit in earlier rounds + in earlier waves of this
//      round + its rank among its peers. Equal digits keep their slot order, so each pass is
//      stable and so is the whole sort.
// Keys below 2^20 (grids up to 1024 x 1024 cells) take three passes.
//
// The claim of dead slots by pending spawns (claim.comp:10-17) uses the same scans, so it
// is defined too: the k-th dead slot in slot order takes the k-th pending particle.
#include "dev.h"

namespace bin {
constexpr int NT = 256;           // work-items per group
constexpr int ITEMS = 16;         // keys per work-item
constexpr int TILE = NT * ITEMS;  // keys per group
constexpr int RADIX = 256;        // 8-bit digits
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_CHUNK = NT * SCAN_ITEMS;  // entries per chunk of the histogram scan
constexpr int TOP_NT = 1024;
constexpr int TOP_ITEMS = 8;
constexpr int TOP_MAX = TOP_NT * TOP_ITEMS;  // most chunks the one-group scan covers
}  // namespace bin

namespace {

// Lanes of this wave with the same 8-bit digit (valid lanes only). Every lane of the wave
// must call it.
SP_DEV uint64_t wave_peers(uint32_t d, bool valid) {
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
        const bool bit = ((d >> b) & 1u) != 0u;
        const uint64_t on = __ballot(bit);
        m &= bit ? on : ~on;
    }
    return m;
}

SP_DEV uint64_t lanes_below() { return (1ull << __lane_id()) - 1ull; }

// Inclusive prefix sum across the 64 lanes of a wave.
SP_DEV uint32_t wave_incl_scan(uint32_t v) {
    const int lane = __lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    return v;
}

// Exclusive prefix sum of one value per work-item over a group of NTG work-items (x only);
// *total receives the group's sum. `lds` holds NTG / 64 words. Every work-item must call it.
template <int NTG>
SP_DEV uint32_t group_excl_scan(uint32_t v, uint32_t* lds, uint32_t* total) {
    const int t = threadIdx.x, w = t >> 6;
    const uint32_t inc = wave_incl_scan(v);
    if ((t & 63) == 63) lds[w] = inc;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < NTG / 64; k++) {
        const uint32_t s = lds[k];
        if (k < w) off += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return off + inc - v;
}

// Key of one particle slot: its cell (cell_of, sort.comp:28-30), or with `tiled` its 8 x 8
// tile and then its cell within the tile; `dead` for a dead slot, larger than every live key.
SP_DEV uint32_t bin_key(const sp_particle& p, int W, int H, int tiled, uint32_t dead) {
    if (p.kind == SP_KIND_DEAD) return dead;
    const int x = clampi_((int)p.pos[0], 0, W - 1), y = clampi_((int)p.pos[1], 0, H - 1);
    if (tiled) return (uint32_t)((((y >> 3) * ((W + 7) >> 3) + (x >> 3)) << 6) | ((y & 7) << 3) | (x & 7));
    return (uint32_t)(y * W + x);
}

}  // namespace

// Keys and slots of caller particles (the handle-free entry points).
__global__ __launch_bounds__(256) void k_bin_keys(const sp_particle* __restrict__ part, uint32_t n, int W, int H, int tiled,
                                                  uint32_t dead, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const uint32_t id = blockIdx.x * 256 + threadIdx.x;
    if (id >= n) return;
    keys[id] = bin_key(part[id], W, H, tiled, dead);
    vals[id] = id;
}

// Keys and slots of a tank's n particle slots.
__global__ __launch_bounds__(256) void k_sort_keys(const TankBatch tb, uint32_t* keys, uint32_t* vals, uint32_t n, int tiled,
                                                   uint32_t dead) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    const uint32_t id = blockIdx.x * 256 + threadIdx.x;
    if (id >= n) return;
    keys[id] = bin_key(c.part[id], c.W, c.H, tiled, dead);
    vals[id] = id;
}

// Pass step 1: digit counts of group b's keys into hist[d * nblk + b].
__global__ __launch_bounds__(bin::NT) void k_bin_hist(const uint32_t* __restrict__ keys, uint32_t n, int shift,
                                                      uint32_t* __restrict__ hist, uint32_t nblk) {
    using namespace bin;
    __shared__ uint32_t cnt[RADIX];
    const int t = threadIdx.x;
    cnt[t] = 0u;
    const uint32_t base = blockIdx.x * (uint32_t)TILE;
    uint32_t k[ITEMS];
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
        const uint32_t i = base + (uint32_t)(r * NT + t);
        k[r] = i < n ? keys[i] : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
        const uint32_t i = base + (uint32_t)(r * NT + t);
        const bool valid = i < n;
        const uint32_t d = (k[r] >> shift) & 255u;
        const uint64_t peers = wave_peers(d, valid);
        if (valid && (peers & lanes_below()) == 0ull) atomicAdd(&cnt[d], (uint32_t)__popcll(peers));
    }
    __syncthreads();
    hist[(size_t)t * nblk + blockIdx.x] = cnt[t];
}

// Pass step 2a: exclusive scan of each SCAN_CHUNK entries of a[0, m) in place; the chunk's
// total into tot[chunk].
__global__ __launch_bounds__(bin::NT) void k_bin_scan_chunks(uint32_t* __restrict__ a, uint32_t m, uint32_t* __restrict__ tot) {
    using namespace bin;
    __shared__ uint32_t lds[NT / 64];
    const uint32_t first = blockIdx.x * (uint32_t)SCAN_CHUNK + threadIdx.x * (uint32_t)SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS], s = 0u;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
        v[j] = first + j < m ? a[first + j] : 0u;
        s += v[j];
    }
    uint32_t total;
    uint32_t run = group_excl_scan<NT>(s, lds, &total);
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
        if (first + j < m) a[first + j] = run;
        run += v[j];
    }
    if (threadIdx.x == 0) tot[blockIdx.x] = total;
}

// Pass step 2b: exclusive scan of the chunks' totals, one group (at most TOP_MAX chunks).
__global__ __launch_bounds__(bin::TOP_NT) void k_bin_scan_top(uint32_t* __restrict__ tot, uint32_t m) {
    using namespace bin;
    __shared__ uint32_t lds[TOP_NT / 64];
    const uint32_t first = threadIdx.x * (uint32_t)TOP_ITEMS;
    uint32_t v[TOP_ITEMS], s = 0u;
#pragma unroll
    for (int j = 0; j < TOP_ITEMS; j++) {
        v[j] = first + j < m ? tot[first + j] : 0u;
        s += v[j];
    }
    uint32_t total;
    uint32_t run = group_excl_scan<TOP_NT>(s, lds, &total);
#pragma unroll
    for (int j = 0; j < TOP_ITEMS; j++) {
        if (first + j < m) tot[first + j] = run;
        run += v[j];
    }
}

// Pass step 3: every key and its slot to its place, in slot order within each digit.
__global__ __launch_bounds__(bin::NT) void k_bin_scatter(const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
                                                         uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out,
                                                         uint32_t n, int shift, const uint32_t* __restrict__ hist,
                                                         const uint32_t* __restrict__ chunk, uint32_t nblk) {
    using namespace bin;
    __shared__ uint32_t start[RADIX];          // where this group's keys of each digit begin
    __shared__ uint32_t wcnt[NT / 64][RADIX];  // keys of each digit per wave in this round
    const int t = threadIdx.x, w = t >> 6;
    {
        const uint32_t e = (uint32_t)t * nblk + blockIdx.x;
        start[t] = hist[e] + chunk[e / (uint32_t)SCAN_CHUNK];
#pragma unroll
        for (int q = 0; q < NT / 64; q++) wcnt[q][t] = 0u;
    }
    const uint32_t base = blockIdx.x * (uint32_t)TILE;
    uint32_t k[ITEMS], v[ITEMS];
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
        const uint32_t i = base + (uint32_t)(r * NT + t);
        k[r] = i < n ? keys_in[i] : 0u;
        v[r] = i < n ? vals_in[i] : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
        if (base + (uint32_t)(r * NT) >= n) break;  // the same for every work-item of the group
        const uint32_t i = base + (uint32_t)(r * NT + t);
        const bool valid = i < n;
        const uint32_t d = (k[r] >> shift) & 255u;
        const uint64_t peers = wave_peers(d, valid);
        const uint32_t rank = (uint32_t)__popcll(peers & lanes_below());
        if (valid && rank == 0u) wcnt[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = start[d] + rank;
            for (int q = 0; q < w; q++) pos += wcnt[q][d];
            keys_out[pos] = k[r];
            vals_out[pos] = v[r];
        }
        __syncthreads();
        uint32_t s = 0u;
#pragma unroll
        for (int q = 0; q < NT / 64; q++) {
            s += wcnt[q][t];
            wcnt[q][t] = 0u;
        }
        start[t] += s;
        __syncthreads();
    }
}

// Particles (and with `sout` their stress) gathered into the sorted order (sort.comp:66).
__global__ __launch_bounds__(256) void k_sort_gather(const TankBatch tb, const uint32_t* vals, sp_particle* out, float4* sout,
                                                     uint32_t n) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    const uint32_t j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const uint32_t from = vals[j];
    out[j] = c.part[from];
    if (sout) sout[j] = c.pstress[from];
}

__global__ __launch_bounds__(256) void k_bin_gather(const sp_particle* __restrict__ in, const uint32_t* __restrict__ vals,
                                                    sp_particle* __restrict__ out, uint32_t* __restrict__ perm, uint32_t n) {
    const uint32_t j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const uint32_t from = vals[j];
    out[j] = in[from];
    if (perm) perm[j] = from;
}

// ---- claim.comp with a defined order --------------------------------------------------

// Dead slots among each group of BIN_TILE of the tank's n slots.
__global__ __launch_bounds__(bin::NT) void k_claim_count(const TankBatch tb, uint32_t n, uint32_t* __restrict__ cnt) {
    using namespace bin;
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    __shared__ uint32_t lds[NT / 64];
    const uint32_t base = blockIdx.x * (uint32_t)TILE;
    uint32_t s = 0u;
    for (int r = 0; r < ITEMS; r++) {
        const uint32_t i = base + (uint32_t)(r * NT) + threadIdx.x;
        if (i < n && c.part[i].kind == SP_KIND_DEAD) s++;
    }
    uint32_t total;
    (void)group_excl_scan<NT>(s, lds, &total);
    if (threadIdx.x == 0) cnt[blockIdx.x] = total;
}

// The k-th dead slot (in slot order) takes spawn[k] while k < spawn_count (claim.comp:10-17).
// off, chunk: the chunk scans (k_bin_scan_chunks, k_bin_scan_top) of k_claim_count's counts.
__global__ __launch_bounds__(bin::NT) void k_claim(const TankBatch tb, uint32_t n, const sp_particle* __restrict__ spawn,
                                                   uint32_t spawn_count, const uint32_t* __restrict__ off,
                                                   const uint32_t* __restrict__ chunk) {
    using namespace bin;
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    __shared__ uint32_t lds[NT / 64];
    const uint32_t base = blockIdx.x * (uint32_t)TILE;
    uint32_t run = off[blockIdx.x] + chunk[blockIdx.x / (uint32_t)SCAN_CHUNK];
    if (run >= spawn_count) return;  // the same for every work-item of the group
    for (int r = 0; r < ITEMS; r++) {
        if (base + (uint32_t)(r * NT) >= n) break;
        const uint32_t i = base + (uint32_t)(r * NT) + threadIdx.x;
        const bool dead = i < n && c.part[i].kind == SP_KIND_DEAD;
        uint32_t total;
        const uint32_t k = run + group_excl_scan<NT>(dead ? 1u : 0u, lds, &total);
        if (dead && k < spawn_count) c.part[i] = spawn[k];
        run += total;
        if (run >= spawn_count) break;  // uniform: every work-item has the same run
    }
}

// ---- host side -------------------------------------------------------------------------

namespace {

// Work space of a sort of n keys: two key and two slot arrays, the histogram and the chunk
// totals, back to back (256-byte aligned).
struct BinSpace {
    uint32_t *keys, *vals, *keys2, *vals2, *hist, *chunk;
};

constexpr uint32_t bin_groups(uint32_t n) {
    return n == 0u ? 1u : (uint32_t)(((uint64_t)n + bin::TILE - 1) / bin::TILE);
}
constexpr uint32_t bin_chunks(uint32_t n) {
    return (uint32_t)(((uint64_t)bin::RADIX * bin_groups(n) + bin::SCAN_CHUNK - 1) / bin::SCAN_CHUNK);
}
constexpr size_t bin_align(size_t b) { return (b + 255) & ~(size_t)255; }

constexpr size_t bin_space_bytes(uint32_t n) {
    return 4 * bin_align(4 * (size_t)(n == 0u ? 1u : n)) + bin_align(4 * (size_t)bin::RADIX * bin_groups(n)) +
           bin_align(4 * (size_t)bin_chunks(n));
}

BinSpace bin_space_at(void* base, uint32_t n) {
    char* p = (char*)base;
    const size_t m = std::max<uint32_t>(n, 1);
    BinSpace s;
    uint32_t** arrays[4] = {&s.keys, &s.vals, &s.keys2, &s.vals2};
    for (uint32_t** a : arrays) {
        *a = (uint32_t*)p;
        p += bin_align(4 * m);
    }
    s.hist = (uint32_t*)p;
    p += bin_align(4 * (size_t)bin::RADIX * bin_groups(n));
    s.chunk = (uint32_t*)p;
    return s;
}

// Bits of the largest live key of a grid (keys below 2^bits; the dead key is 2^bits - 1).
int bin_bits(int W, int H, int tiled) {
    const uint32_t live_max = tiled ? (uint32_t)(((W + 7) / 8) * ((H + 7) / 8) * 64) : (uint32_t)(W * H);
    int bits = 1;
    while (bits < 32 && (1u << bits) <= live_max) bits++;
    return bits;
}
uint32_t bin_dead_key(int bits) { return bits >= 32 ? 0xffffffffu : (1u << bits) - 1u; }

// The radix passes over the n (key, slot) pairs in sp.keys / sp.vals; returns the slot
// array in sorted order (sp.vals or sp.vals2).
const uint32_t* launch_bin_passes(BinSpace& sp, uint32_t n, int bits, hipStream_t stream) {
    const uint32_t nblk = bin_groups(n), m = bin::RADIX * nblk, nch = bin_chunks(n);
    uint32_t *ki = sp.keys, *vi = sp.vals, *ko = sp.keys2, *vo = sp.vals2;
    for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(k_bin_hist, dim3(nblk), dim3(bin::NT), 0, stream, (const uint32_t*)ki, n, shift, sp.hist, nblk);
        hipLaunchKernelGGL(k_bin_scan_chunks, dim3(nch), dim3(bin::NT), 0, stream, sp.hist, m, sp.chunk);
        hipLaunchKernelGGL(k_bin_scan_top, dim3(1), dim3(bin::TOP_NT), 0, stream, sp.chunk, nch);
        hipLaunchKernelGGL(k_bin_scatter, dim3(nblk), dim3(bin::NT), 0, stream, (const uint32_t*)ki, (const uint32_t*)vi, ko, vo, n,
                           shift, (const uint32_t*)sp.hist, (const uint32_t*)sp.chunk, nblk);
        std::swap(ki, ko);
        std::swap(vi, vo);
    }
    return vi;
}

}  // namespace
