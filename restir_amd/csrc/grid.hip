// World-space hash grid passes on gfx950: clear, and insert in two modes.
//
// Grid buffer layout (hashgrid.sh:32-33, 52-59): uint32
//   [0, T)            checksums (0: empty slot)
//   [T, 2T)           counts (may exceed the 32 listed)
//   [2T, 2T + 32T)    entries, 32 per slot
//
// Atomic mode (hashgrid.sh:182-200): each wave covers an 8 x 8 pixel tile, so
// the pixels of a cell mostly share a wave. Lanes with identical keys follow
// identical probe sequences, so one leader lane per distinct key runs the
// compare-and-swap probing for all of them and takes all their entry indices
// with one atomicAdd: a legal ordering of the per-pixel atomics of the
// reference, with far fewer L2 atomics.
//
// Deterministic mode: the result of one fixed linearization of those atomics.
// Round p = 0..7 of probing handles the pixels still unresolved, in ascending
// pixel index: a pixel claims its probe slot if empty, joins it if it holds
// its checksum, else moves to round p + 1. On the GPU a round is a claim step
// (atomicMin of (round tag, pixel index) on empty target slots) and a resolve
// step (the minimum claimant wins; others join iff their checksum equals the
// winner's). Round 0 runs over all pixels; the few pixels left run rounds
// 1..7 in one workgroup. The adds are then ordered by pixel index: each cell's
// count is every pixel resolved to it and its entries are its 32 lowest pixel
// indices, ascending, found by a CSR gather and a wave bitonic sort per cell.
#include "device.h"
#include "batch.h"
#include "internal.h"

namespace rgi {
namespace {

__global__ void __launch_bounds__(256) k_grid_clear(uint32_t* grid, uint32_t words)
{
    // 2T words from a 256-byte aligned base: whole uint4 stores.
    uint32_t vec = words >> 2;
    uint4* g4 = reinterpret_cast<uint4*>(grid);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < vec; i += gridDim.x * blockDim.x)
        g4[i] = make_uint4(0u, 0u, 0u, 0u);
}

RGI_DEV uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }
RGI_DEV uint32_t rank_in(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// 16 x 16 pixels per block (block bx, by of the grid), one 8 x 8 tile per
// wave.
RGI_DEV void block_pixel(uint32_t bx, uint32_t by, int& x, int& y)
{
    uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    x = (int)(bx * 16u + (wave & 1u) * 8u + (lane & 7u));
    y = (int)(by * 16u + (wave >> 1) * 8u + (lane >> 3));
}
RGI_DEV void block_pixel(int& x, int& y) { block_pixel(blockIdx.x, blockIdx.y, x, y); }

// gridInsert with its atomics aggregated per wave, for the workgroup at
// (bx, by) of the 16 x 16-pixel grid.
RGI_DEV void grid_insert_body(const rgi_frame& f, const rgi_gbuffer& g, F2 texel, uint32_t* grid, uint32_t bx,
                              uint32_t by)
{
    int x, y;
    block_pixel(bx, by, x, y);
    uint32_t table = f.grid_table_size, mask = table - 1u;
    uint32_t slot0 = 0, checksum = 0;
    bool active = x < f.gi_width && y < f.gi_height && insert_key(f, g, texel, x, y, slot0, checksum);
    uint32_t pixel = (uint32_t)y * (uint32_t)f.gi_width + (uint32_t)x;
    uint32_t lane = lane_id();
    uint64_t pending = __ballot(active);
    while (pending) {
        uint32_t leader = (uint32_t)__builtin_ctzll(pending);
        uint32_t lead_slot = __builtin_amdgcn_readlane(slot0, leader);
        uint32_t lead_sum = __builtin_amdgcn_readlane(checksum, leader);
        bool peer = ((pending >> lane) & 1ull) && slot0 == lead_slot && checksum == lead_sum;
        uint64_t peers = __ballot(peer);
        uint32_t resolved = kGridInvalid, base = 0;
        if (lane == leader) {
            for (uint32_t probe = 0; probe < kGridMaxProbes; ++probe) {
                uint32_t s = (lead_slot + probe) & mask;
                uint32_t previous = atomicCAS(&grid[s], 0u, lead_sum);
                if (previous == 0u || previous == lead_sum) {
                    resolved = s;
                    break;
                }
            }
            if (resolved != kGridInvalid) base = atomicAdd(&grid[table + resolved], (uint32_t)__popcll(peers));
        }
        resolved = __builtin_amdgcn_readlane(resolved, leader);
        base = __builtin_amdgcn_readlane(base, leader);
        if (peer && resolved != kGridInvalid) {
            uint32_t index = base + rank_in(peers);
            if (index < kGridCellCapacity) grid[2u * table + resolved * kGridCellCapacity + index] = pixel;
        }
        pending &= ~peers;
    }
}

__global__ void __launch_bounds__(256) k_grid_insert_atomic(rgi_frame f, rgi_gbuffer g, F2 texel, uint32_t* grid)
{
    grid_insert_body(f, g, texel, grid, blockIdx.x, blockIdx.y);
}

// The batched forms: workgroups [first[v], first[v + 1]) work on view v.
__global__ void __launch_bounds__(256) k_grid_clear_batch(const BatchView* views, const uint32_t* first, uint32_t count)
{
    uint32_t b;
    const uint32_t vi = batch_view(first, count, blockIdx.x, b);
    const BatchView& v = views[vi];
    const uint32_t blocks = first[vi + 1u] - first[vi];
    const uint32_t vec = (2u * v.f.grid_table_size) >> 2;
    uint4* g4 = reinterpret_cast<uint4*>(v.grid);
    for (uint32_t i = b * blockDim.x + threadIdx.x; i < vec; i += blocks * blockDim.x) g4[i] = make_uint4(0u, 0u, 0u, 0u);
}

__global__ void __launch_bounds__(256) k_grid_insert_batch(const BatchView* views, const uint32_t* first, uint32_t count)
{
    uint32_t b;
    const BatchView& v = views[batch_view(first, count, blockIdx.x, b)];
    grid_insert_body(v.f, v.g, v.gi_texel, v.grid, b % v.grid_gx, b / v.grid_gx);
}

// ------------------------------------------------------- deterministic mode

struct DetScratch {
    uint32_t* slot0;     // [n] first probe slot
    uint32_t* checksum;  // [n] 0: no insert
    uint32_t* resolved;  // [n] slot, or invalid
    uint32_t* pending;   // [n] pixels left after round 0
    uint32_t* pending2;  // [n] ping-pong of the tail rounds
    uint32_t* list;      // [n] CSR pixel lists
    uint32_t* claim;     // [T] claim tags during probing; slot -> cell index afterwards
    uint32_t* occupied;  // [min(T, n)] slots claimed, in claim order
    uint32_t* offset;    // [min(T, n)] CSR offsets per cell
    uint32_t* fill;      // [min(T, n)] CSR fill counters per cell
    uint32_t* counters;  // [0] pending count, [1] occupied count, [2] resolved count
    uint32_t* block_sum; // [min(T, n) / 1024 + 1] listed pixels per 1024 occupied cells, then their offsets
};

constexpr uint32_t kScanCells = 1024; // occupied cells per workgroup of the offsets scan

constexpr uint32_t kTagShift = 29;
constexpr uint32_t kIndexMask = (1u << kTagShift) - 1u;
RGI_DEV uint32_t claim_tag(uint32_t round, uint32_t pixel) { return ((7u - round) << kTagShift) | pixel; }

__global__ void __launch_bounds__(256) k_det_keys(rgi_frame f, rgi_gbuffer g, F2 texel, DetScratch d)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t n = (uint32_t)f.gi_width * (uint32_t)f.gi_height;
    if (i >= n) return;
    int x = (int)(i % (uint32_t)f.gi_width), y = (int)(i / (uint32_t)f.gi_width);
    uint32_t slot = 0, checksum = 0;
    if (!insert_key(f, g, texel, x, y, slot, checksum)) checksum = 0;
    d.slot0[i] = slot;
    d.checksum[i] = checksum;
    d.resolved[i] = kGridInvalid;
}

__global__ void __launch_bounds__(256) k_fill(uint32_t* p, uint32_t n, uint32_t value)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = value;
}

__global__ void __launch_bounds__(256) k_det_claim0(uint32_t n, uint32_t mask, const uint32_t* grid, DetScratch d)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || d.checksum[i] == 0u) return;
    uint32_t s = d.slot0[i] & mask;
    if (grid[s] == 0u) atomicMin(&d.claim[s], claim_tag(0u, i));
}

// Resolution of pixel i in round p without its counter updates: true when it
// found its slot `s`; `claimed` when it is the slot's new owner (it writes
// the slot's checksum).
RGI_DEV bool det_resolve_core(uint32_t i, uint32_t round, uint32_t table, uint32_t* grid, DetScratch& d,
                              uint32_t& s, bool& claimed)
{
    uint32_t mask = table - 1u;
    uint32_t cs = d.checksum[i];
    s = (d.slot0[i] + round) & mask;
    uint32_t c = d.claim[s];
    bool done;
    claimed = false;
    if (c != 0xffffffffu && (c >> kTagShift) == 7u - round) {
        // Empty at the start of this round: the lowest claimant took it.
        uint32_t winner = c & kIndexMask;
        done = d.checksum[winner] == cs;
        if (winner == i) {
            grid[s] = cs;
            claimed = true;
        }
    } else {
        done = grid[s] == cs;
    }
    if (done) d.resolved[i] = s;
    return done;
}

// The same with one atomic per counter update (the tail's lanes diverge).
RGI_DEV bool det_resolve(uint32_t i, uint32_t round, uint32_t table, uint32_t* grid, DetScratch& d)
{
    uint32_t s;
    bool claimed;
    bool done = det_resolve_core(i, round, table, grid, d, s, claimed);
    if (claimed) d.occupied[atomicAdd(&d.counters[1], 1u)] = s;
    if (done) {
        atomicAdd(&grid[table + s], 1u);
        atomicAdd(&d.counters[2], 1u);
    }
    return done;
}

// Wave-aggregated fetch-and-add of 1 per active lane to *addr (addresses may
// differ between lanes): one atomicAdd per distinct address in the wave, each
// lane receiving the old value plus its rank among the lanes sharing its
// address. All lanes of the wave must call it. The sums do not depend on the
// order of the atomics, and every consumer of the returned positions (the
// occupied-cell and pending lists, a cell's gathered pixels) is order-free:
// pending pixels claim by atomicMin, CSR offsets follow whatever order the
// cells come in, and each cell's pixels are sorted afterwards.
RGI_DEV uint32_t aggregated_add(bool active, uint32_t* addr)
{
    const uint32_t lane = lane_id();
    const uint64_t a = reinterpret_cast<uint64_t>(addr);
    uint64_t pending = __ballot(active);
    uint32_t result = 0;
    while (pending) {
        uint32_t leader = (uint32_t)__builtin_ctzll(pending);
        // readlane returns an int: widen each half through uint32_t, never
        // by sign extension.
        uint32_t lead_hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(a >> 32), (int)leader);
        uint32_t lead_lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)a, (int)leader);
        uint64_t lead = ((uint64_t)lead_hi << 32) | (uint64_t)lead_lo;
        // The leader is always its own peer, so every pass retires at least
        // one lane and the loop ends after at most 64 passes.
        bool peer = lane == leader || (((pending >> lane) & 1ull) && a == lead);
        uint64_t peers = __ballot(peer);
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(addr, (uint32_t)__popcll(peers));
        base = __builtin_amdgcn_readlane(base, leader);
        if (peer) result = base + rank_in(peers);
        pending &= ~peers;
    }
    return result;
}

// Pixel index of this thread over the GI pixels (gw x gh) in 16 x 16 blocks
// of 8 x 8 tiles per wave, so that a wave's pixels share few cells; false
// past the edge.
RGI_DEV bool tile_index(int gw, int gh, uint32_t& i)
{
    int x, y;
    block_pixel(x, y);
    i = (uint32_t)y * (uint32_t)gw + (uint32_t)x;
    return x < gw && y < gh;
}

// Positions for the lanes of this workgroup with `flag` set in a list whose
// length is *counter: each wave's lanes are ranked by ballot, the waves by an
// LDS counter, the workgroup by one global atomicAdd (not one per wave: a
// global counter that every wave increments serializes in L2). All threads
// of the workgroup must call it.
RGI_DEV uint32_t workgroup_append(bool flag, uint32_t* counter, uint32_t* lds_count, uint32_t* lds_base)
{
    const uint32_t lane = lane_id();
    if (threadIdx.x == 0) *lds_count = 0u;
    __syncthreads();
    uint64_t m = __ballot(flag);
    uint32_t wave_base = 0;
    if (lane == 0u && m != 0ull) wave_base = atomicAdd(lds_count, (uint32_t)__popcll(m));
    wave_base = __builtin_amdgcn_readlane(wave_base, 0);
    __syncthreads();
    if (threadIdx.x == 0) *lds_base = *lds_count ? atomicAdd(counter, *lds_count) : 0u;
    __syncthreads();
    return *lds_base + wave_base + rank_in(m);
}

__global__ void __launch_bounds__(256) k_det_resolve0(int gw, int gh, uint32_t table, uint32_t* grid, DetScratch d)
{
    __shared__ uint32_t lds_count, lds_base;
    uint32_t i;
    bool active = tile_index(gw, gh, i);
    active = active && d.checksum[i] != 0u;
    uint32_t s = 0;
    bool claimed = false, done = false;
    if (active) done = det_resolve_core(i, 0u, table, grid, d, s, claimed);
    uint32_t k = workgroup_append(claimed, &d.counters[1], &lds_count, &lds_base);
    if (claimed) d.occupied[k] = s;
    aggregated_add(done, &grid[table + s]);
    workgroup_append(done, &d.counters[2], &lds_count, &lds_base);
    bool left = active && !done;
    uint32_t p = workgroup_append(left, &d.counters[0], &lds_count, &lds_base);
    if (left) d.pending[p] = i;
}

// Rounds 1..7 over the pixels round 0 left, in one workgroup: a block
// barrier separates each claim step from its resolve step.
__global__ void __launch_bounds__(1024) k_det_tail(uint32_t table, uint32_t* grid, DetScratch d)
{
    __shared__ uint32_t next_count;
    uint32_t mask = table - 1u;
    uint32_t count = d.counters[0];
    uint32_t* cur = d.pending;
    uint32_t* nxt = d.pending2;
    for (uint32_t round = 1; round < kGridMaxProbes && count > 0u; ++round) {
        if (threadIdx.x == 0) next_count = 0u;
        for (uint32_t k = threadIdx.x; k < count; k += blockDim.x) {
            uint32_t i = cur[k];
            uint32_t s = (d.slot0[i] + round) & mask;
            if (grid[s] == 0u) atomicMin(&d.claim[s], claim_tag(round, i));
        }
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < count; k += blockDim.x) {
            uint32_t i = cur[k];
            if (!det_resolve(i, round, table, grid, d)) nxt[atomicAdd(&next_count, 1u)] = i;
        }
        __syncthreads();
        count = next_count;
        uint32_t* t = cur;
        cur = nxt;
        nxt = t;
        __syncthreads();
    }
}

// CSR offsets of the occupied cells (cells in claim order), a three-step
// scan: per workgroup of kScanCells cells the sum of their counts; one
// workgroup's exclusive scan over those sums; per workgroup the exclusive
// scan of its cells' counts plus its sum's offset. Each cell's offset is the
// pixels listed before it, its fill counter starts at 0, and its slot maps to
// its cell index.
RGI_DEV uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds, uint32_t& total)
{
    // 256 threads: Hillis-Steele over LDS.
    uint32_t t = threadIdx.x;
    lds[t] = v;
    __syncthreads();
    for (uint32_t stride = 1; stride < 256u; stride <<= 1) {
        uint32_t add = t >= stride ? lds[t - stride] : 0u;
        __syncthreads();
        lds[t] += add;
        __syncthreads();
    }
    total = lds[255];
    uint32_t inclusive = lds[t];
    __syncthreads();
    return inclusive - v;
}

__global__ void __launch_bounds__(256) k_det_block_sums(uint32_t table, const uint32_t* grid, DetScratch d)
{
    __shared__ uint32_t lds[256];
    const uint32_t cells = d.counters[1];
    const uint32_t first = blockIdx.x * kScanCells;
    if (first >= cells) return; // the same for the whole workgroup
    uint32_t sum = 0;
    for (uint32_t k = first + threadIdx.x; k < first + kScanCells && k < cells; k += 256u)
        sum += grid[table + d.occupied[k]];
    uint32_t total;
    block_exclusive_scan(sum, lds, total);
    if (threadIdx.x == 0) d.block_sum[blockIdx.x] = total;
}

__global__ void __launch_bounds__(256) k_det_block_scan(DetScratch d)
{
    __shared__ uint32_t lds[256];
    const uint32_t blocks = (d.counters[1] + kScanCells - 1u) / kScanCells;
    // Up to 16 block sums per thread, in order.
    const uint32_t per = (blocks + 255u) / 256u;
    const uint32_t begin = threadIdx.x * per;
    uint32_t sum = 0;
    for (uint32_t b = begin; b < begin + per && b < blocks; ++b) sum += d.block_sum[b];
    uint32_t total;
    uint32_t running = block_exclusive_scan(sum, lds, total);
    for (uint32_t b = begin; b < begin + per && b < blocks; ++b) {
        uint32_t v = d.block_sum[b];
        d.block_sum[b] = running;
        running += v;
    }
}

__global__ void __launch_bounds__(256) k_det_offsets(uint32_t table, const uint32_t* grid, DetScratch d)
{
    __shared__ uint32_t lds[256];
    const uint32_t cells = d.counters[1];
    const uint32_t first = blockIdx.x * kScanCells;
    if (first >= cells) return; // the same for the whole workgroup
    // Thread t owns cells first + 4t .. first + 4t + 3.
    constexpr uint32_t kPer = kScanCells / 256u;
    uint32_t counts[kPer], sum = 0;
#pragma unroll
    for (uint32_t j = 0; j < kPer; ++j) {
        uint32_t k = first + threadIdx.x * kPer + j;
        counts[j] = k < cells ? grid[table + d.occupied[k]] : 0u;
        sum += counts[j];
    }
    uint32_t total;
    uint32_t running = d.block_sum[blockIdx.x] + block_exclusive_scan(sum, lds, total);
#pragma unroll
    for (uint32_t j = 0; j < kPer; ++j) {
        uint32_t k = first + threadIdx.x * kPer + j;
        if (k < cells) {
            d.offset[k] = running;
            d.fill[k] = 0u;
            d.claim[d.occupied[k]] = k;
        }
        running += counts[j];
    }
}

__global__ void __launch_bounds__(256) k_det_scatter(int gw, int gh, DetScratch d)
{
    uint32_t i;
    bool inside = tile_index(gw, gh, i);
    uint32_t s = inside ? d.resolved[i] : kGridInvalid;
    bool active = s != kGridInvalid;
    uint32_t k = active ? d.claim[s] : 0u;
    uint32_t rank = aggregated_add(active, &d.fill[k]);
    if (active) {
        uint32_t pos = d.offset[k] + rank;
        if (pos < d.counters[2]) d.list[pos] = i;
    }
}

RGI_DEV uint32_t bitonic_sort64(uint32_t v, uint32_t lane)
{
    for (uint32_t k = 2; k <= 64; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            uint32_t other = (uint32_t)__shfl_xor((int)v, (int)j, 64);
            bool ascending = (lane & k) == 0u;
            bool lower = (lane & j) == 0u;
            uint32_t lo = v < other ? v : other, hi = v < other ? other : v;
            v = (lower == ascending) ? lo : hi;
        }
    return v;
}

// One wave per occupied cell: its 32 lowest pixel indices, ascending.
__global__ void __launch_bounds__(256) k_det_select(uint32_t table, uint32_t* grid, DetScratch d)
{
    uint32_t lane = threadIdx.x & 63u;
    uint32_t cells = d.counters[1];
    for (uint32_t k = blockIdx.x * 4u + (threadIdx.x >> 6); k < cells; k += gridDim.x * 4u) {
        uint32_t s = d.occupied[k];
        // Counts the grid held before this insert are not in the lists: read
        // no further than what this insert listed.
        uint32_t n = grid[table + s];
        uint32_t listed = d.counters[2] - d.offset[k];
        n = n < listed ? n : listed;
        const uint32_t* src = d.list + d.offset[k];
        uint32_t v = lane < n ? src[lane] : 0xffffffffu;
        v = bitonic_sort64(v, lane);
        for (uint32_t c = 64; c < n; c += 32) {
            if (lane >= 32u) v = c + (lane - 32u) < n ? src[c + (lane - 32u)] : 0xffffffffu;
            v = bitonic_sort64(v, lane);
        }
        if (lane < kGridCellCapacity && lane < n) grid[2u * table + s * kGridCellCapacity + lane] = v;
    }
}

} // namespace

int grid_clear(uint32_t* grid, uint32_t table_size, hipStream_t stream)
{
    uint32_t words = 2u * table_size;
    hipLaunchKernelGGL(k_grid_clear, dim3(grid_clear_blocks(table_size)), dim3(256), 0, stream, grid, words);
    return check_launch("grid clear");
}

// u_screenInfo.zw: the GI pixels' texel size; the G-buffer is read at its own size.
static F2 texel_of(const rgi_frame& f) { return F2{1.0f / (float)f.gi_width, 1.0f / (float)f.gi_height}; }

int grid_insert_atomic(const rgi_frame& f, const rgi_gbuffer& g, uint32_t* grid, hipStream_t stream)
{
    dim3 blocks((uint32_t)(f.gi_width + 15) / 16u, (uint32_t)(f.gi_height + 15) / 16u);
    hipLaunchKernelGGL(k_grid_insert_atomic, blocks, dim3(256), 0, stream, f, g, texel_of(f), grid);
    return check_launch("grid insert");
}

size_t det_scratch_bytes(uint32_t pixels, uint32_t table)
{
    uint32_t cells = table < pixels ? table : pixels;
    return align256((size_t)pixels * 4) * 6 + align256((size_t)table * 4) + align256((size_t)cells * 4) * 3 + 256 +
           align256((size_t)(cells / kScanCells + 1) * 4);
}

int grid_insert_deterministic(rgi_context* ctx, const rgi_frame& f, const rgi_gbuffer& g, uint32_t* grid,
                              hipStream_t stream)
{
    uint32_t n = (uint32_t)f.gi_width * (uint32_t)f.gi_height;
    char* base = static_cast<char*>(context_scratch(ctx, det_scratch_bytes(n, f.grid_table_size)));
    if (!base) return RGI_E_NOMEM;
    return grid_insert_deterministic_in(base, f, g, grid, stream);
}

int grid_insert_deterministic_in(void* scratch_base, const rgi_frame& f, const rgi_gbuffer& g, uint32_t* grid,
                                 hipStream_t stream)
{
    uint32_t n = (uint32_t)f.gi_width * (uint32_t)f.gi_height;
    uint32_t table = f.grid_table_size;
    uint32_t cells = table < n ? table : n;
    char* base = static_cast<char*>(scratch_base);
    DetScratch d;
    auto take = [&](size_t count) {
        uint32_t* p = reinterpret_cast<uint32_t*>(base);
        base += align256(count * 4);
        return p;
    };
    d.slot0 = take(n);
    d.checksum = take(n);
    d.resolved = take(n);
    d.pending = take(n);
    d.pending2 = take(n);
    d.list = take(n);
    d.claim = take(table);
    d.occupied = take(cells);
    d.offset = take(cells);
    d.fill = take(cells);
    d.counters = take(64);
    d.block_sum = take(cells / kScanCells + 1);
    uint32_t blocks = (n + 255u) / 256u;
    hipLaunchKernelGGL(k_det_keys, dim3(blocks), dim3(256), 0, stream, f, g, texel_of(f), d);
    hipLaunchKernelGGL(k_fill, dim3(1024), dim3(256), 0, stream, d.claim, table, 0xffffffffu);
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(256), 0, stream, d.counters, 64u, 0u);
    hipLaunchKernelGGL(k_det_claim0, dim3(blocks), dim3(256), 0, stream, n, table - 1u, grid, d);
    const dim3 tiles((uint32_t)(f.gi_width + 15) / 16u, (uint32_t)(f.gi_height + 15) / 16u);
    hipLaunchKernelGGL(k_det_resolve0, tiles, dim3(256), 0, stream, f.gi_width, f.gi_height, table, grid, d);
    hipLaunchKernelGGL(k_det_tail, dim3(1), dim3(1024), 0, stream, table, grid, d);
    const uint32_t scan_blocks = cells / kScanCells + 1; // covers every occupied cell (at most `cells`)
    hipLaunchKernelGGL(k_det_block_sums, dim3(scan_blocks), dim3(256), 0, stream, table, grid, d);
    hipLaunchKernelGGL(k_det_block_scan, dim3(1), dim3(256), 0, stream, d);
    hipLaunchKernelGGL(k_det_offsets, dim3(scan_blocks), dim3(256), 0, stream, table, grid, d);
    hipLaunchKernelGGL(k_det_scatter, tiles, dim3(256), 0, stream, f.gi_width, f.gi_height, d);
    hipLaunchKernelGGL(k_det_select, dim3(2048), dim3(256), 0, stream, table, grid, d);
    return check_launch("grid insert (deterministic)");
}

uint32_t grid_clear_blocks(uint32_t table_size)
{
    uint32_t blocks = (2u * table_size / 4u + 255u) / 256u;
    return blocks > 4096u ? 4096u : (blocks == 0 ? 1u : blocks);
}

int grid_clear_batch(const BatchView* views, const uint32_t* first, uint32_t count, uint32_t blocks, hipStream_t stream)
{
    if (blocks == 0u) return RGI_OK;
    hipLaunchKernelGGL(k_grid_clear_batch, dim3(blocks), dim3(256), 0, stream, views, first, count);
    return check_launch("grid clear (batch)");
}

int grid_insert_batch(const BatchView* views, const uint32_t* first, uint32_t count, uint32_t blocks,
                      hipStream_t stream)
{
    if (blocks == 0u) return RGI_OK;
    hipLaunchKernelGGL(k_grid_insert_batch, dim3(blocks), dim3(256), 0, stream, views, first, count);
    return check_launch("grid insert (batch)");
}

} // namespace rgi
