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

// 16 x 16 pixels per block, one 8 x 8 tile per wave.
RGI_DEV void block_pixel(int& x, int& y)
{
    uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    x = (int)(blockIdx.x * 16u + (wave & 1u) * 8u + (lane & 7u));
    y = (int)(blockIdx.y * 16u + (wave >> 1) * 8u + (lane >> 3));
}

__global__ void __launch_bounds__(256) k_grid_insert_atomic(rgi_frame f, rgi_gbuffer g, F2 texel, uint32_t* grid)
{
    int x, y;
    block_pixel(x, y);
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
};

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

// Resolution of pixel i in round p; true when it found its slot.
RGI_DEV bool det_resolve(uint32_t i, uint32_t round, uint32_t table, uint32_t* grid, DetScratch& d)
{
    uint32_t mask = table - 1u;
    uint32_t cs = d.checksum[i];
    uint32_t s = (d.slot0[i] + round) & mask;
    uint32_t c = d.claim[s];
    bool done;
    if (c != 0xffffffffu && (c >> kTagShift) == 7u - round) {
        // Empty at the start of this round: the lowest claimant took it.
        uint32_t winner = c & kIndexMask;
        done = d.checksum[winner] == cs;
        if (winner == i) {
            grid[s] = cs;
            uint32_t k = atomicAdd(&d.counters[1], 1u);
            d.occupied[k] = s;
        }
    } else {
        done = grid[s] == cs;
    }
    if (done) {
        d.resolved[i] = s;
        atomicAdd(&grid[table + s], 1u);
        atomicAdd(&d.counters[2], 1u);
    }
    return done;
}

__global__ void __launch_bounds__(256) k_det_resolve0(uint32_t n, uint32_t table, uint32_t* grid, DetScratch d)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || d.checksum[i] == 0u) return;
    if (!det_resolve(i, 0u, table, grid, d)) {
        uint32_t k = atomicAdd(&d.counters[0], 1u);
        d.pending[k] = i;
    }
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

// CSR offsets of the occupied cells (one workgroup; cells in claim order).
__global__ void __launch_bounds__(1024) k_det_offsets(uint32_t table, const uint32_t* grid, DetScratch d)
{
    __shared__ uint32_t partial[1024];
    uint32_t cells = d.counters[1];
    uint32_t per = (cells + blockDim.x - 1u) / blockDim.x;
    uint32_t begin = threadIdx.x * per;
    uint32_t end = begin + per < cells ? begin + per : cells;
    uint32_t sum = 0;
    for (uint32_t k = begin; k < end; ++k) sum += grid[table + d.occupied[k]];
    partial[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t stride = 1; stride < blockDim.x; stride <<= 1) {
        uint32_t v = threadIdx.x >= stride ? partial[threadIdx.x - stride] : 0u;
        __syncthreads();
        partial[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t running = partial[threadIdx.x] - sum;
    for (uint32_t k = begin; k < end; ++k) {
        uint32_t s = d.occupied[k];
        d.offset[k] = running;
        d.fill[k] = 0u;
        d.claim[s] = k;
        running += grid[table + s];
    }
}

__global__ void __launch_bounds__(256) k_det_scatter(uint32_t n, DetScratch d)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t s = d.resolved[i];
    if (s == kGridInvalid) return;
    uint32_t k = d.claim[s];
    uint32_t pos = d.offset[k] + atomicAdd(&d.fill[k], 1u);
    if (pos < d.counters[2]) d.list[pos] = i;
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
    uint32_t blocks = (words / 4u + 255u) / 256u;
    blocks = blocks > 4096u ? 4096u : (blocks == 0 ? 1u : blocks);
    hipLaunchKernelGGL(k_grid_clear, dim3(blocks), dim3(256), 0, stream, grid, words);
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
    return align256((size_t)pixels * 4) * 6 + align256((size_t)table * 4) + align256((size_t)cells * 4) * 3 + 256;
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
    uint32_t blocks = (n + 255u) / 256u;
    hipLaunchKernelGGL(k_det_keys, dim3(blocks), dim3(256), 0, stream, f, g, texel_of(f), d);
    hipLaunchKernelGGL(k_fill, dim3(1024), dim3(256), 0, stream, d.claim, table, 0xffffffffu);
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(256), 0, stream, d.counters, 64u, 0u);
    hipLaunchKernelGGL(k_det_claim0, dim3(blocks), dim3(256), 0, stream, n, table - 1u, grid, d);
    hipLaunchKernelGGL(k_det_resolve0, dim3(blocks), dim3(256), 0, stream, n, table, grid, d);
    hipLaunchKernelGGL(k_det_tail, dim3(1), dim3(1024), 0, stream, table, grid, d);
    hipLaunchKernelGGL(k_det_offsets, dim3(1), dim3(1024), 0, stream, table, grid, d);
    hipLaunchKernelGGL(k_det_scatter, dim3(blocks), dim3(256), 0, stream, n, d);
    hipLaunchKernelGGL(k_det_select, dim3(2048), dim3(256), 0, stream, table, grid, d);
    return check_launch("grid insert (deterministic)");
}

} // namespace rgi
