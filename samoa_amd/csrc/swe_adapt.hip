// swe_adapt.hip -- remeshing on the device: the reference's adaption step
// (SWE.f90:387-390) for a mesh held in forward curve order.
//
//   1. conformity (Conformity/Conformity.f90:172-208): every cell-edge holds a
//      2-bit word (bit 0 refine, bit 1 coarsen); an edge's value is the OR of
//      its two refine bits and the AND of its two coarsen bits, which is what
//      the reference's single shared edge accumulates.  The requests set the
//      words (mark_edges, :1244-1268); sweeps of mark_edges_integrity
//      (:1271-1317) run until no word changes.  All updates are monotone, so
//      the fixpoint does not depend on the order of the sweeps.
//   2. tree position: cell i covers [off_i, off_i + 2^(60 - depth_i)) of the
//      two roots [0, 2^61); a first child starts on a multiple of its
//      parent's size.  off is an exclusive scan of the cell sizes.
//   3. counts of new cells (split 2..4, merge 1 for the first child of a
//      pair and 0 for the second, else 1), an exclusive scan of them.
//   4. emit: depth, plotter type, tree offset and state of every new cell (a
//      split copies eta, hu, hv and b, or averages b over the child from a
//      raster as get_bathymetry_at_element does under _ADAPT_INTEGRATE; a
//      merge averages all four, SWE_adapt.f90:95-169); the neighbours inside
//      each old cell, and between two cells that keep their place; for a
//      changing cell, the new cell-edge on each half of each old edge.  One
//      kernel streams over the kept cells, another over the changing ones.
//   5. link, over the changing cells only: a new edge on half s of old edge k
//      of cell A meets the new edge on the same half of the same old edge in
//      A's neighbour B; the same end of the shared edge is half s in B when A
//      and B run in opposite senses (the plotter types differ in sign) and
//      half 1-s when they run in the same.  Both sides are written.
//
// Cell corners are (entry, right angle, exit); a cell (e, r, x) splits into
// (e, m, r) and (r, m, x), m the midpoint of the hypotenuse, with the plotter
// types child(p), -child(-p) of SFC_node_traversal.f90:257,270.  Edge k is
// opposite corner k; edge 1 is the hypotenuse; half 0 of edge k is the half
// at corner (k+1) mod 3, half 1 the half at corner (k+2) mod 3.
#include <hip/hip_runtime.h>

#include <chrono>

#include "swe_internal.hpp"
#include "swe_sweep_stamp.hpp"

namespace swe {
namespace {

constexpr int AT = 256;           // threads per workgroup of the per-cell kernels
// the emit kernels' workgroups, measured: the kept-cell emit streams over
// every cell (larger workgroups, fewer of them), the changing-cell emit and
// link run over the ~5 % that change (smaller ones, spread over more CUs)
constexpr int KEEP_T = 512;
constexpr int CHANGE_T = 128;
constexpr int SCAN_T = 256;       // scan: threads per tile
constexpr int SCAN_ITEMS = 8;     // scan: consecutive items per thread
constexpr int SCAN_TILE = SCAN_T * SCAN_ITEMS;
constexpr int TOP_T = 1024;       // scan of the tile sums: one workgroup

// child(p) of SFC_node_traversal.f90:257,270.  The map is an involution, so
// the same table gives the parent type of a first child: child(child(p)) = p.
__constant__ int8_t c_child[17] = {3, 2, 1, 8, 7, 6, 5, 4, 0, -6, -7, -8, -1, -2, -3, -4, -5};

// New cells of a split: pattern 0 hypotenuse only, 1 with leg 2 (the first
// child splits again), 2 with leg 0 (the second child splits again), 3 both.
// Per new cell: depth increment, type rule (0: child(p), 1: -child(-p),
// 2/3: the same rules applied to the first child's type, 4/5: to the second
// child's) and per edge either 0x10 | (old edge << 2) | half (half 2: the
// whole old edge) or 0x20 | (cell of the pattern << 2) | its edge.  Each new
// cell is one 64-bit word (byte 0 the increment, byte 1 the rule, bytes 2-4
// the edges), so that a pattern is read with two 16-B loads.
constexpr uint64_t split_cell(int dd, int rule, int e0, int e1, int e2) {
    return (uint64_t)(uint8_t)dd | ((uint64_t)(uint8_t)rule << 8) | ((uint64_t)(uint8_t)e0 << 16) |
           ((uint64_t)(uint8_t)e1 << 24) | ((uint64_t)(uint8_t)e2 << 32);
}
__constant__ __align__(16) uint64_t c_split[4][4] = {
    {split_cell(1, 0, 0x20 | (1 << 2) | 2, 0x10 | (2 << 2) | 2, 0x10 | (1 << 2) | 1),
     split_cell(1, 1, 0x10 | (1 << 2) | 0, 0x10 | (0 << 2) | 2, 0x20 | (0 << 2) | 0), 0, 0},
    {split_cell(2, 2, 0x20 | (1 << 2) | 2, 0x10 | (1 << 2) | 1, 0x10 | (2 << 2) | 0),
     split_cell(2, 3, 0x10 | (2 << 2) | 1, 0x20 | (2 << 2) | 2, 0x20 | (0 << 2) | 0),
     split_cell(1, 1, 0x10 | (1 << 2) | 0, 0x10 | (0 << 2) | 2, 0x20 | (1 << 2) | 1), 0},
    {split_cell(1, 0, 0x20 | (1 << 2) | 1, 0x10 | (2 << 2) | 2, 0x10 | (1 << 2) | 1),
     split_cell(2, 4, 0x20 | (2 << 2) | 2, 0x20 | (0 << 2) | 0, 0x10 | (0 << 2) | 0),
     split_cell(2, 5, 0x10 | (0 << 2) | 1, 0x10 | (1 << 2) | 0, 0x20 | (1 << 2) | 0), 0},
    {split_cell(2, 2, 0x20 | (1 << 2) | 2, 0x10 | (1 << 2) | 1, 0x10 | (2 << 2) | 0),
     split_cell(2, 3, 0x10 | (2 << 2) | 1, 0x20 | (2 << 2) | 1, 0x20 | (0 << 2) | 0),
     split_cell(2, 4, 0x20 | (3 << 2) | 2, 0x20 | (1 << 2) | 1, 0x10 | (0 << 2) | 0),
     split_cell(2, 5, 0x10 | (0 << 2) | 1, 0x10 | (1 << 2) | 0, 0x20 | (2 << 2) | 0)},
};

// codes of the count kernel
constexpr int8_t CODE_KEEP = 0, CODE_MERGE_FIRST = 1, CODE_MERGE_SECOND = 2, CODE_SPLIT = 4;

__device__ __forceinline__ uint32_t ld_relaxed(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Value (bit 0 refine, bit 1 coarsen) of edge k of cell c.
__device__ __forceinline__ uint32_t edge_value(const uint32_t *slots, const int32_t *nbr, int64_t c,
                                               uint32_t own, int k) {
    const uint32_t s = (own >> (2 * k)) & 3u;
    const int32_t q = nbr[3 * c + k];
    if (q < 0) return s;
    const uint32_t t = (ld_relaxed(&slots[q >> 2]) >> (2 * (q & 3))) & 3u;
    return ((s | t) & 1u) | (s & t & 2u);
}

// mark_edges (Conformity.f90:1244-1268): the word a request f gives a cell
// of depth d.  The hypotenuse is never coarsened, a cell that keeps its
// depth clears its legs' coarsen bits, a refinement request marks the
// hypotenuse.  Requests beyond [dmin, dmax] are dropped as volume_op
// (SWE_euler_timestep.f90:339-342) never issues them.
__device__ __forceinline__ uint32_t request_word(int f, int d, int dmin, int dmax) {
    if (f > 0 && d >= dmax) f = 0;
    if (f < 0 && d <= dmin) f = 0;
    const uint32_t legs = f != 0 ? 2u : 0u;
    const uint32_t hyp = f > 0 ? 1u : 0u;
    return legs | (hyp << 2) | (legs << 4);
}

// The word the integrity rule gives a cell from the values (bit 0 refine,
// bit 1 coarsen) of its three edges: a refined edge refines the hypotenuse,
// and the legs stay coarsened only if both are and nothing is refined.
__device__ __forceinline__ uint32_t integrity_word(uint32_t v0, uint32_t v1, uint32_t v2, int d) {
    const uint32_t rh = (v0 | v1 | v2) & 1u;
    const uint32_t cc = (!rh && (v0 & v2 & 2u) && d > 0) ? 2u : 0u;
    return cc | (rh << 2) | (cc << 4);
}

// mark_edges_integrity (Conformity.f90:1271-1317), in place, over tiles of
// SWEEP_TILE cells, SWEEP_CPT per thread (the sweep is bound by each
// workgroup's chain of memory round trips: more cells per workgroup, fewer
// workgroups in turn).  A workgroup iterates its tile to a fixpoint of its
// own in LDS (a neighbour in the tile is read there, any other from memory
// every round) before it writes the tile back, so chains of forced
// refinements inside a tile settle in one sweep.  Every update is monotone,
// which keeps the global fixpoint the same.
//
// The first sweep of a remeshing starts from the requests: every word, its
// own and an outside neighbour's, is request_word of the flags, an earlier
// state than any other tile may have written; it writes every word of its
// tile.
//
// Sweep numbers carry the call's tag in their high 16 bits (sweep_number).
// stamp[T] is the last sweep in which a cell next to tile T but outside it
// changed (or T itself did not settle).  From the second sweep on, a tile
// whose stamp is older than the previous sweep (a stamp of an earlier call
// included, compared modulo 2^32: stamp_is_old) has read every neighbour word
// since it last changed and is at its fixpoint: it is skipped.  *changed gets the sweep if a word changed.
constexpr int SWEEP_CPT = 2;
constexpr int SWEEP_TILE = AT * SWEEP_CPT;
__global__ __launch_bounds__(AT) void conf_sweep_kernel(int64_t n, const int8_t *depth,
                                                        const int32_t *nbr, uint32_t *slots,
                                                        uint32_t *stamp, uint32_t *changed,
                                                        uint32_t sweep, const int8_t *flags,
                                                        int dmin, int dmax) {
    constexpr int CPT = SWEEP_CPT, TILE = SWEEP_TILE;
    __shared__ uint32_t w[TILE];
    __shared__ uint32_t s_skip;
    const bool first = (sweep & 0xffffu) == 1u;
    if (!first) {
        if (threadIdx.x == 0) s_skip = stamp_is_old(ld_relaxed(&stamp[blockIdx.x]), sweep) ? 1u : 0u;
        __syncthreads();
        if (s_skip) return;  // uniform across the workgroup
    }
    const int64_t t0 = blockIdx.x * (int64_t)TILE;
    // cell j of the thread is t0 + j AT + threadIdx.x, so that each load is
    // coalesced across the workgroup
    int64_t c[CPT];
    bool live[CPT];
    int32_t q[CPT][3];
    int d[CPT];
    uint32_t own[CPT], start[CPT];
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        c[j] = t0 + j * AT + threadIdx.x;
        live[j] = c[j] < n;
        const int64_t cc = live[j] ? c[j] : 0;
        q[j][0] = live[j] ? nbr[3 * cc] : -1;
        q[j][1] = live[j] ? nbr[3 * cc + 1] : -1;
        q[j][2] = live[j] ? nbr[3 * cc + 2] : -1;
        d[j] = live[j] ? depth[cc] : 0;
        own[j] = !live[j] ? 0u : first ? request_word(flags[cc], d[j], dmin, dmax)
                                       : ld_relaxed(&slots[cc]);
    }
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        start[j] = own[j];
        w[j * AT + threadIdx.x] = own[j];
    }
    // a neighbour outside the tile is read from memory: at its own index
    // (oi), every lane at some index (its own cell, for a neighbour in the
    // tile or none), so that the loads are issued together.  In the first
    // sweep those words are the requests, fixed for the whole sweep, and are
    // read once.
    bool in_tile[CPT][3];
    int64_t oi[CPT][3];
    uint32_t fw[CPT][3];
#pragma unroll
    for (int j = 0; j < CPT; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int64_t o = q[j][k] >> 2;
            in_tile[j][k] = q[j][k] < 0 || (o >= t0 && o < t0 + TILE);
            oi[j][k] = in_tile[j][k] ? (live[j] ? c[j] : 0) : o;
            fw[j][k] = 0u;
        }
    if (first) {
        int8_t fo[CPT][3], dp[CPT][3];
#pragma unroll
        for (int j = 0; j < CPT; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                fo[j][k] = flags[oi[j][k]];
                dp[j][k] = depth[oi[j][k]];
            }
#pragma unroll
        for (int j = 0; j < CPT; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k) fw[j][k] = request_word(fo[j][k], dp[j][k], dmin, dmax);
    }
    __syncthreads();
    // each round settles at least one more cell of a chain; a tile that has
    // not settled after 32 rounds stamps itself for the next sweep
    bool settled = false;
    for (int round = 0; round < 32; ++round) {
        uint32_t ow[CPT][3];
#pragma unroll
        for (int j = 0; j < CPT; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k) ow[j][k] = fw[j][k];
        if (!first) {
            uint32_t sl[CPT][3];
#pragma unroll
            for (int j = 0; j < CPT; ++j)
#pragma unroll
                for (int k = 0; k < 3; ++k) sl[j][k] = ld_relaxed(&slots[oi[j][k]]);
#pragma unroll
            for (int j = 0; j < CPT; ++j)
#pragma unroll
                for (int k = 0; k < 3; ++k) ow[j][k] = sl[j][k];
        }
        uint32_t nw[CPT];
        bool moved = false;
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            nw[j] = own[j];
            if (live[j]) {
                uint32_t v[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const uint32_t sk = (own[j] >> (2 * k)) & 3u;
                    v[k] = sk;
                    if (q[j][k] >= 0) {
                        const uint32_t wk =
                            in_tile[j][k] ? w[(q[j][k] >> 2) - t0] : ow[j][k];
                        const uint32_t tk = (wk >> (2 * (q[j][k] & 3))) & 3u;
                        v[k] = ((sk | tk) & 1u) | (sk & tk & 2u);
                    }
                }
                nw[j] = integrity_word(v[0], v[1], v[2], d[j]);
            }
            moved |= nw[j] != own[j];
        }
        __syncthreads();  // every read of this round is done
#pragma unroll
        for (int j = 0; j < CPT; ++j)
            if (nw[j] != own[j]) {
                w[j * AT + threadIdx.x] = nw[j];
                own[j] = nw[j];
            }
        if (!__syncthreads_or(moved)) {
            settled = true;
            break;
        }
    }
    bool any = false;
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const bool diff = own[j] != start[j];
        any |= diff;
        if (live[j] && (diff || first)) st_relaxed(&slots[c[j]], own[j]);
        if (diff) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                if (q[j][k] < 0) continue;
                const int64_t o = q[j][k] >> 2;
                if (o < t0 || o >= t0 + TILE) {
                    uint32_t *st = &stamp[o / TILE];
                    if (ld_relaxed(st) != sweep) st_relaxed(st, sweep);
                }
            }
        }
    }
    if (!settled && threadIdx.x == 0) st_relaxed(&stamp[blockIdx.x], sweep);
    // one status write per workgroup, and only if no other has written it:
    // stores from every wave to one word contend across the whole device
    if (__syncthreads_or(any) && threadIdx.x == 0 && ld_relaxed(changed) != sweep)
        st_relaxed(changed, sweep);
}

__device__ __forceinline__ int64_t wave_incl_scan(int64_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int64_t u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    return v;
}

// Exclusive scan over the workgroup (NT threads); *total gets the sum.
template <int NT>
__device__ int64_t block_excl_scan(int64_t v, int64_t *lds, int64_t *total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t inc = wave_incl_scan(v);
    if (lane == 63) lds[wave] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t run = 0;
        for (int w = 0; w < NT / 64; ++w) {
            const int64_t t = lds[w];
            lds[w] = run;
            run += t;
        }
        lds[NT / 64] = run;
    }
    __syncthreads();
    const int64_t out = lds[wave] + inc - v;
    *total = lds[NT / 64];
    __syncthreads();
    return out;
}

__global__ __launch_bounds__(SCAN_T) void scan_reduce_kernel(const int64_t *in, int64_t n,
                                                             int64_t *bsum) {
    __shared__ int64_t lds[SCAN_T / 64 + 1];
    const int64_t base = blockIdx.x * (int64_t)SCAN_TILE + threadIdx.x * (int64_t)SCAN_ITEMS;
    int64_t s = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j)
        if (base + j < n) s += in[base + j];
    int64_t total;
    (void)block_excl_scan<SCAN_T>(s, lds, &total);
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// Exclusive scan of the nb tile sums in place, by one workgroup of TOP_T
// threads; returns the grand total.
__device__ int64_t scan_tile_sums(int64_t *bsum, int64_t nb) {
    __shared__ int64_t lds[TOP_T / 64 + 1];
    int64_t carry = 0;
    for (int64_t start = 0; start < nb; start += TOP_T) {
        const int64_t i = start + threadIdx.x;
        const int64_t v = i < nb ? bsum[i] : 0;
        int64_t total;
        const int64_t ex = block_excl_scan<TOP_T>(v, lds, &total);
        if (i < nb) bsum[i] = carry + ex;
        carry += total;
    }
    return carry;
}

// bsum[0..nb) scanned in place; bsum[nb] = the grand total.
__global__ __launch_bounds__(TOP_T) void scan_top_kernel(int64_t *bsum, int64_t nb) {
    const int64_t total = scan_tile_sums(bsum, nb);
    if (threadIdx.x == 0) bsum[nb] = total;
}

// The scan of the count pass's tile sums, then its results straight into
// host memory (coherent, mapped): out[0] = res[0] (the covered part of the
// root trees, when this call scanned the offsets), out[1] = the new cell
// count, out[2] = status bits (0: an inconsistent merge, 1: the mesh
// changes, 2: not at the fixpoint), out[3] = the last sweep that changed a
// word, out[5] = the cells that change.  out[7] = seq goes last, a release at
// system scope: a host that reads seq there reads the rest too.  Clears the
// error bits for the next count pass.
__global__ __launch_bounds__(TOP_T) void plan_top_kernel(int64_t *bsum, int64_t nb,
                                                         uint32_t *status, const int64_t *res,
                                                         int64_t *out, int64_t seq) {
    const int64_t total = scan_tile_sums(bsum, nb);
    if (threadIdx.x == 0) {
        bsum[nb] = total;
        out[0] = res[0];
        out[1] = total & 0xffffffffLL;
        out[2] = (int64_t)(status[0] | ((total >> 32) != 0 ? 2u : 0u));
        out[3] = (int64_t)status[1];
        out[5] = total >> 32;
        status[0] = 0u;
        __hip_atomic_store(&out[7], seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// out[i] = exclusive prefix of in (in place allowed: each thread reads its
// items before it writes them).
__global__ __launch_bounds__(SCAN_T) void scan_apply_kernel(const int64_t *in, int64_t *out,
                                                            int64_t n, const int64_t *bsum) {
    __shared__ int64_t lds[SCAN_T / 64 + 1];
    const int64_t base = blockIdx.x * (int64_t)SCAN_TILE + threadIdx.x * (int64_t)SCAN_ITEMS;
    int64_t v[SCAN_ITEMS];
    int64_t s = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        v[j] = base + j < n ? in[base + j] : 0;
        s += v[j];
    }
    int64_t total;
    int64_t run = bsum[blockIdx.x] + block_excl_scan<SCAN_T>(s, lds, &total);
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j)
        if (base + j < n) {
            out[base + j] = run;
            run += v[j];
        }
}

__global__ __launch_bounds__(AT) void tree_size_kernel(int64_t n, const int8_t *depth,
                                                       int64_t *size) {
    const int64_t c = blockIdx.x * (int64_t)AT + threadIdx.x;
    if (c < n) size[c] = (int64_t)1 << (60 - depth[c]);
}

// New cells of an old cell from its decision code, with the changing-cell
// count in the high word (a split makes 2..4 cells, the first of a merging
// pair 1, the second 0).
__device__ __forceinline__ int64_t code_count(int k) {
    const int64_t m = k >= CODE_SPLIT ? 2 + (k & 1) + ((k >> 1) & 1) : (k == CODE_MERGE_SECOND ? 0 : 1);
    return m | ((int64_t)(k != CODE_KEEP) << 32);
}

// Decision of every cell, and the sum of the new-cell counts of each scan
// tile.  It also checks the fixpoint the sweeps reached: status bit 2 if the
// integrity rule would still change a word (more sweeps are needed, and
// nothing else here is valid), bit 0 if a coarsening cell's sibling is not a
// coarsening leaf (cannot happen at the fixpoint).  A sibling in the tile is
// checked from LDS, where every cell leaves its depth and the two edge values
// that check reads; one outside it from memory.
__global__ __launch_bounds__(SCAN_T) void count_reduce_kernel(int64_t n, const int8_t *depth,
                                                              const int32_t *nbr,
                                                              const uint32_t *slots,
                                                              const int64_t *off, int8_t *code,
                                                              int64_t *bsum, uint32_t *status) {
    __shared__ int64_t lds[SCAN_T / 64 + 1];
    // per cell: depth << 2 | edge 1 refines << 1 | edge 0 coarsens
    __shared__ uint8_t s_info[SCAN_TILE];
    const int64_t tile = blockIdx.x * (int64_t)SCAN_TILE;
    int64_t sum = 0;
    bool moved = false, bad = false;
    // the sweeps are over: the words are read with plain loads, and every
    // cell's loads are issued before any is used (own word, depth and
    // neighbours; then the neighbours' words and, where the cell's own word
    // could coarsen edge 0, its tree offset), each at an index in range
    uint32_t own[SCAN_ITEMS], nw[SCAN_ITEMS][3];
    int32_t q[SCAN_ITEMS][3];
    int dep[SCAN_ITEMS];
    int64_t ofs[SCAN_ITEMS];
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        const int64_t c = tile + j * SCAN_T + threadIdx.x;
        const bool live = c < n;
        const int64_t cc = live ? c : 0;
        own[j] = slots[cc];
        dep[j] = depth[cc];
        const int32_t q0 = nbr[3 * cc], q1 = nbr[3 * cc + 1], q2 = nbr[3 * cc + 2];
        q[j][0] = live ? q0 : -1;
        q[j][1] = live ? q1 : -1;
        q[j][2] = live ? q2 : -1;
    }
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        const int64_t c = tile + j * SCAN_T + threadIdx.x;
        const int64_t cc = c < n ? c : 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) nw[j][k] = slots[q[j][k] >= 0 ? q[j][k] >> 2 : cc];
        ofs[j] = (own[j] & 2u) ? off[cc] : 0;
    }
    uint32_t vv[SCAN_ITEMS];
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        const int64_t c = tile + j * SCAN_T + threadIdx.x;
        uint32_t v[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const uint32_t sk = (own[j] >> (2 * k)) & 3u;
            const uint32_t tk = (nw[j][k] >> (2 * (q[j][k] & 3))) & 3u;
            v[k] = q[j][k] < 0 ? sk : (((sk | tk) & 1u) | (sk & tk & 2u));
        }
        vv[j] = v[0] | (v[1] << 2) | (v[2] << 4);
        if (c < n) {
            moved |= integrity_word(v[0], v[1], v[2], dep[j]) != own[j];
            s_info[j * SCAN_T + threadIdx.x] =
                (uint8_t)((dep[j] << 2) | ((v[1] & 1u) << 1) | ((v[0] >> 1) & 1u));
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        const int64_t c = tile + j * SCAN_T + threadIdx.x;
        if (c >= n) continue;
        const uint32_t v0 = vv[j] & 3u, v1 = (vv[j] >> 2) & 3u, v2 = (vv[j] >> 4) & 3u;
        const int d = dep[j];
        int k = CODE_KEEP;
        if (v1 & 1u) {
            k = CODE_SPLIT + ((int)(v2 & 1u) | ((int)(v0 & 1u) << 1));
        } else if (v0 & 2u) {
            // d > 0 here: the integrity rule clears the coarsen bits of root cells
            const int64_t parent = (int64_t)1 << (61 - d);
            const bool first = ofs[j] % parent == 0;
            const int64_t sib = first ? c + 1 : c - 1;
            bool ok;
            if (sib >= tile && sib < tile + SCAN_TILE && sib < n) {
                const uint32_t inf = s_info[sib - tile];
                ok = (int)(inf >> 2) == d && (inf & 1u) && !(inf & 2u);
            } else {
                ok = sib >= 0 && sib < n && depth[sib] == d;
                if (ok) {
                    const uint32_t so = slots[sib];
                    ok = (edge_value(slots, nbr, sib, so, 0) & 2u) &&
                         !(edge_value(slots, nbr, sib, so, 1) & 1u);
                }
            }
            bad |= !ok;
            k = first ? CODE_MERGE_FIRST : CODE_MERGE_SECOND;
        }
        code[c] = (int8_t)k;
        sum += code_count(k);
    }
    int64_t total;
    (void)block_excl_scan<SCAN_T>(sum, lds, &total);
    const uint32_t err = (__syncthreads_or(moved) ? 4u : 0u) | (__syncthreads_or(bad) ? 1u : 0u);
    if (threadIdx.x == 0) {
        bsum[blockIdx.x] = total;
        if (err) atomicOr(status, err);
    }
}

// base[c] = exclusive prefix of the counts, rebuilt from the decision codes
// (8 consecutive cells per thread, bsum holding the scanned tile sums), for
// the cells that change only; kidx[c] = the new index of a cell that keeps
// its place (its base), -1 for one that changes; list = the changing cells
// in curve order.  The prefixes go through LDS so that the stores are
// coalesced.
__global__ __launch_bounds__(SCAN_T) void scan_codes_kernel(int64_t n, const int8_t *code,
                                                            const int64_t *bsum, int64_t *base,
                                                            int32_t *kidx, int32_t *list) {
    __shared__ int64_t lds[SCAN_T / 64 + 1];
    __shared__ int64_t stage[SCAN_TILE + SCAN_TILE / SCAN_ITEMS];  // item i at i + i / 8
    __shared__ uint8_t kstage[SCAN_TILE + SCAN_TILE / SCAN_ITEMS];  // 1: the cell is kept
    const int64_t tile = blockIdx.x * (int64_t)SCAN_TILE;
    const int64_t first = tile + threadIdx.x * (int64_t)SCAN_ITEMS;
    int64_t v[SCAN_ITEMS];
    int64_t sum = 0;
    if (first + SCAN_ITEMS <= n) {
        const uint64_t packed = *reinterpret_cast<const uint64_t *>(code + first);
#pragma unroll
        for (int j = 0; j < SCAN_ITEMS; ++j) {
            v[j] = code_count((int)(int8_t)(packed >> (8 * j)));
            sum += v[j];
        }
    } else {
#pragma unroll
        for (int j = 0; j < SCAN_ITEMS; ++j) {
            v[j] = first + j < n ? code_count(code[first + j]) : 0;
            sum += v[j];
        }
    }
    int64_t total;
    int64_t run = bsum[blockIdx.x] + block_excl_scan<SCAN_T>(sum, lds, &total);
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        stage[threadIdx.x * (SCAN_ITEMS + 1) + j] = run;
        kstage[threadIdx.x * (SCAN_ITEMS + 1) + j] = v[j] == 1 ? 1 : 0;
        // the changing cells, in order: the high word counts those before
        if (v[j] >> 32) list[run >> 32] = (int32_t)(first + j);
        run += v[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        const int i = j * SCAN_T + threadIdx.x;
        if (tile + i < n) {
            const int64_t b = stage[i + i / SCAN_ITEMS];
            // a kept cell counts exactly 1 (one new cell, no change); its
            // new index is below 2^29, and -1 marks the others
            const bool kept = kstage[i + i / SCAN_ITEMS] != 0;
            kidx[tile + i] = kept ? (int32_t)(b & 0xffffffffLL) : -1;
            if (!kept) base[tile + i] = b;
        }
    }
}

// Corners (unit square; entry, right angle, exit) of the cell at tree offset
// `off` and depth d, from the corners c of its ancestor at depth d0 (or of
// its root, d0 = 0): one halving per level along the path the offset spells
// (bit 60 - l chooses the second child at level l), with the host's
// arithmetic, m = 0.5 (x1 + x3).  Starting from an ancestor runs the same
// operations as starting from the root.
__device__ void tree_corners(int64_t off, int d0, int d, double c[6]) {
    if (d0 == 0) {
        if ((off >> 60) == 0) {
            c[0] = 0.0; c[1] = 0.0; c[2] = 1.0; c[3] = 0.0; c[4] = 1.0; c[5] = 1.0;
        } else {
            c[0] = 1.0; c[1] = 1.0; c[2] = 0.0; c[3] = 1.0; c[4] = 0.0; c[5] = 0.0;
        }
    }
    double ax = c[0], ay = c[1], bx = c[2], by = c[3], ex = c[4], ey = c[5];
    for (int l = d0 + 1; l <= d; ++l) {
        const double mx = 0.5 * (ax + ex), my = 0.5 * (ay + ey);
        if ((off >> (60 - l)) & 1) {  // (x2, m, x3)
            ax = bx;
            ay = by;
        } else {  // (x1, m, x2)
            ex = bx;
            ey = by;
        }
        bx = mx;
        by = my;
    }
    c[0] = ax; c[1] = ay; c[2] = bx; c[3] = by; c[4] = ex; c[5] = ey;
}

// Bilinear sample of the raster at the world point (X, Y), clamped to its
// last cell (the host's scenario code evaluates the same expression); a point
// outside the nodes' extent gets -5000 m, the sea floor the reference assumes
// outside its data (SWE_initialize.f90:164-171).
__device__ double raster_sample(const AdaptArgs &a, double X, double Y) {
    if (!(a.rx0 <= X && a.ry0 <= Y && X <= a.rx0 + (double)(a.rnx - 1) * a.rdx &&
          Y <= a.ry0 + (double)(a.rny - 1) * a.rdy))
        return -5000.0;
    const double fx = (X - a.rx0) / a.rdx, fy = (Y - a.ry0) / a.rdy;
    const int64_t ix = min(max((int64_t)fx, (int64_t)0), a.rnx - 2);
    const int64_t iy = min(max((int64_t)fy, (int64_t)0), a.rny - 2);
    const double wx = fx - (double)ix, wy = fy - (double)iy;
    const double *r0 = a.raster + iy * a.rnx + ix, *r1 = r0 + a.rnx;
    return (1.0 - wx) * (1.0 - wy) * r0[0] + wx * (1.0 - wy) * r0[1] + (1.0 - wx) * wy * r1[0] +
           wx * wy * r1[1];
}

// b of a new cell of corners c and depth d from the raster, as
// get_bathymetry_at_element computes it under _ADAPT_INTEGRATE
// (SWE_initialize.f90:106-147) for gridded data: L = min(16, raster_cap - d)
// halvings of the cell (raster_cap = min(dmax + 3, data_depth + 3)), each
// level 0.5 (value of the first half (x1, m, x2) + value of the second
// (x2, m, x3)), the raster at the centroid scaling * (x1 + x2 + x3) / 3 +
// offset at the bottom; L <= 0 is that centroid sample alone.  The 2^L
// leaves run in depth-first order, a leaf's corners found by L halvings
// along its index; a stack of partial values combines each finished pair as
// the recursion does.
__device__ double raster_cell(const AdaptArgs &a, const double c[6], int d) {
    const int levels = min(16, a.raster_cap - d);
    if (levels <= 0) {  // a cell finer than the data: its centroid (no stack)
        const double cx = (c[0] + c[2] + c[4]) / 3.0, cy = (c[1] + c[3] + c[5]) / 3.0;
        return raster_sample(a, a.scaling * cx + a.offset_x, a.scaling * cy + a.offset_y);
    }
    double stack[17];
    int sp = 0;
    const int leaves = levels > 0 ? 1 << levels : 1;
    for (int i = 0; i < leaves; ++i) {
        double ax = c[0], ay = c[1], bx = c[2], by = c[3], ex = c[4], ey = c[5];
        for (int l = levels - 1; l >= 0; --l) {
            const double mx = 0.5 * (ax + ex), my = 0.5 * (ay + ey);
            if ((i >> l) & 1) {  // (x2, m, x3)
                ax = bx;
                ay = by;
            } else {  // (x1, m, x2)
                ex = bx;
                ey = by;
            }
            bx = mx;
            by = my;
        }
        const double cx = (ax + bx + ex) / 3.0, cy = (ay + by + ey) / 3.0;
        double v = raster_sample(a, a.scaling * cx + a.offset_x, a.scaling * cy + a.offset_y);
        for (int j = i; j & 1; j >>= 1) v = 0.5 * (stack[--sp] + v);
        stack[sp++] = v;
    }
    return stack[0];
}

__device__ __forceinline__ void put_slot(int32_t *eslot, int64_t c, int k, int half, int32_t v) {
    if (half == 2) {
        eslot[6 * c + 2 * k] = v;
        eslot[6 * c + 2 * k + 1] = v;
    } else {
        eslot[6 * c + 2 * k + half] = v;
    }
}

// The new cell-edge on half s of old edge k of cell c: computed for a cell
// that keeps its place, read from the slots emit_change_kernel wrote otherwise.
__device__ __forceinline__ int32_t edge_slot(const AdaptArgs &a, int64_t c, int k, int s) {
    const int32_t kc = a.kidx[c];
    if (kc >= 0) return (kc << 2) | k;
    return a.eslot[6 * c + 2 * k + s];
}

// The cells that keep their place: a shifted copy of depth, plotter type,
// tree offset and state, and their neighbours.  An edge to a cell that keeps
// its place too keeps its pairing; an edge to a changing cell gets -1 here and
// its pairing from link_kernel, which writes both sides of every edge of the
// changing cells.  Writing all three entries makes the neighbour store one
// 12-B store per lane, dense across the wave.
__global__ __launch_bounds__(KEEP_T) void emit_keep_kernel(AdaptArgs a) {
    const int64_t c = blockIdx.x * (int64_t)KEEP_T + threadIdx.x;
    if (c >= a.n) return;
    if (c == 0 && a.scen_dev) *a.scen_dev = a.scen_new;  // no host upload of the table
    // every load, the neighbours' new indices included, is issued before the
    // test (a changing cell, one in twenty, reads them for nothing); the
    // empty asm keeps the compiler from sinking them below it, which would
    // cost a memory round trip each
    const int32_t kc = a.kidx[c];
    const int8_t dc = a.depth[c], pc = a.plotter[c];
    const int64_t oc = a.off[c];
    const double ec = a.eta[c], uc = a.hu[c], vc = a.hv[c], bc = a.b[c];
    const int32_t q0 = a.nbr[3 * c], q1 = a.nbr[3 * c + 1], q2 = a.nbr[3 * c + 2];
    const int32_t k0 = a.kidx[q0 >= 0 ? q0 >> 2 : c], k1 = a.kidx[q1 >= 0 ? q1 >> 2 : c],
                  k2 = a.kidx[q2 >= 0 ? q2 >> 2 : c];
    asm volatile("" ::"v"((int)dc), "v"((int)pc), "v"(oc), "v"(ec), "v"(uc), "v"(vc), "v"(bc), "v"(k0),
                 "v"(k1), "v"(k2));
    if (kc < 0) return;  // a changing cell: emit_change_kernel
    const int64_t b0 = kc;
    a.depth_out[b0] = dc;
    a.plotter_out[b0] = pc;
    a.off_out[b0] = oc;
    a.eta_out[b0] = ec;
    a.hu_out[b0] = uc;
    a.hv_out[b0] = vc;
    a.b_out[b0] = bc;
    a.nbr_out[3 * b0] = q0 >= 0 && k0 >= 0 ? (k0 << 2) | (q0 & 3) : -1;
    a.nbr_out[3 * b0 + 1] = q1 >= 0 && k1 >= 0 ? (k1 << 2) | (q1 & 3) : -1;
    a.nbr_out[3 * b0 + 2] = q2 >= 0 && k2 >= 0 ? (k2 << 2) | (q2 & 3) : -1;
}

// The cells that split or merge (list[i], in curve order): their new cells,
// and for every half of every old edge the new cell-edge that lies on it.
// The loads run in four rounds, each issued at once: the list entry; the
// cell (and, for a merge, its sibling c + 1); its split pattern and the
// children's types; the grandchildren's types.  The state stays in registers: as far as the
// compiler knows, the new cells' stores could alias the old arrays, and it
// would read them again after every store.
__global__ __launch_bounds__(CHANGE_T) void emit_change_kernel(AdaptArgs a) {
    const int64_t i = blockIdx.x * (int64_t)CHANGE_T + threadIdx.x;
    if (i >= a.n_changing) return;
    const int64_t c = a.list[i];
    const int64_t c1 = c + 1 < a.n ? c + 1 : c;
    const int d = a.depth[c];
    const int p = a.plotter[c];
    const int k = a.code[c];
    const int64_t b0 = a.base[c] & 0xffffffffLL;  // the high word counts changing cells
    const int64_t off0 = a.off[c];
    const double ec = a.eta[c], uc = a.hu[c], vc = a.hv[c], bc = a.b[c];
    const double e1 = a.eta[c1], u1 = a.hu[c1], v1 = a.hv[c1], b1 = a.b[c1];
    asm volatile("" ::"v"(d), "v"(p), "v"(k), "v"(b0), "v"(off0), "v"(ec), "v"(uc), "v"(vc),
                 "v"(bc), "v"(e1), "v"(u1), "v"(v1), "v"(b1));
    if (k == CODE_MERGE_FIRST) {
        a.depth_out[b0] = (int8_t)(d - 1);
        a.plotter_out[b0] = c_child[p + 8];  // the parent's type
        a.off_out[b0] = off0;
        a.eta_out[b0] = 0.5 * (ec + e1);
        a.hu_out[b0] = 0.5 * (uc + u1);
        a.hv_out[b0] = 0.5 * (vc + v1);
        a.b_out[b0] = 0.5 * (bc + b1);
        // the first child's hypotenuse is the parent's leg 2, its outer leg
        // half of the parent's hypotenuse; its leg 0 disappears
        put_slot(a.eslot, c, 1, 2, (int32_t)((b0 << 2) | 2));
        put_slot(a.eslot, c, 2, 2, (int32_t)((b0 << 2) | 1));
        put_slot(a.eslot, c, 0, 2, -1);
    } else if (k == CODE_MERGE_SECOND) {
        const int64_t par = b0 - 1;  // emitted by the first child
        put_slot(a.eslot, c, 1, 2, (int32_t)((par << 2) | 0));
        put_slot(a.eslot, c, 0, 2, (int32_t)((par << 2) | 1));
        put_slot(a.eslot, c, 2, 2, -1);
    } else {
        const int pat = k - CODE_SPLIT;
        const int cells = 2 + (pat & 1) + (pat >> 1);
        const uint64_t w0 = c_split[pat][0], w1 = c_split[pat][1], w2 = c_split[pat][2],
                       w3 = c_split[pat][3];
        // the types of the rules: child(p), -child(-p), then both applied to
        // each of those two
        const int t0 = c_child[p + 8], t1 = -c_child[-p + 8];
        const int t2 = c_child[t0 + 8], t3 = -c_child[-t0 + 8], t4 = c_child[t1 + 8],
                  t5 = -c_child[-t1 + 8];
        int64_t coff = off0;  // tree offset of the next child
        double pc[6];         // the parent's corners, for the raster
        if (a.raster) tree_corners(coff, 0, d, pc);
        for (int j = 0; j < cells; ++j) {
            const uint64_t w = j == 0 ? w0 : j == 1 ? w1 : j == 2 ? w2 : w3;
            const int rule = (int)((w >> 8) & 0xff);
            const int64_t idx = b0 + j;
            const int dc = d + (int)(w & 0xff);
            a.depth_out[idx] = (int8_t)dc;
            a.plotter_out[idx] = (int8_t)(rule == 0   ? t0
                                          : rule == 1 ? t1
                                          : rule == 2 ? t2
                                          : rule == 3 ? t3
                                          : rule == 4 ? t4
                                                      : t5);
            a.off_out[idx] = coff;
            a.eta_out[idx] = ec;
            a.hu_out[idx] = uc;
            a.hv_out[idx] = vc;
            if (a.raster) {
                double cc[6] = {pc[0], pc[1], pc[2], pc[3], pc[4], pc[5]};
                tree_corners(coff, d, dc, cc);
                a.b_out[idx] = raster_cell(a, cc, dc);
            } else {
                a.b_out[idx] = bc;
            }
            coff += (int64_t)1 << (60 - dc);
            for (int e = 0; e < 3; ++e) {
                const int src = (int)((w >> (16 + 8 * e)) & 0xff);
                if (src & 0x20) {
                    a.nbr_out[3 * idx + e] = (int32_t)(((b0 + ((src >> 2) & 3)) << 2) | (src & 3));
                } else {
                    put_slot(a.eslot, c, (src >> 2) & 3, src & 3, (int32_t)((idx << 2) | e));
                }
            }
        }
    }
}

// Over the changing cells: each new edge on an old edge meets the new edge on
// the same half across it, and both sides are written (the other side may
// keep its place and then leaves the pairing to this kernel).  The loads run
// in four rounds, each issued at once: the cell; its slots, neighbours and
// type; the neighbours' types and new indices; the slots of the neighbours
// that change (read at every lane, at an index in range, and used where the
// neighbour changes).
__global__ __launch_bounds__(CHANGE_T) void link_kernel(AdaptArgs a) {
    const int64_t i = blockIdx.x * (int64_t)CHANGE_T + threadIdx.x;
    if (i >= a.n_changing) return;
    const int64_t c = a.list[i];
    const int pc = a.plotter[c];
    const int2 *es = reinterpret_cast<const int2 *>(a.eslot + 6 * c);  // 24 B, 8-B aligned
    const int2 x01 = es[0], x23 = es[1], x45 = es[2];
    const int32_t x[3][2] = {{x01.x, x01.y}, {x23.x, x23.y}, {x45.x, x45.y}};
    const int32_t q[3] = {a.nbr[3 * c], a.nbr[3 * c + 1], a.nbr[3 * c + 2]};
    int64_t o[3];
    int ko[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        o[k] = q[k] >= 0 ? q[k] >> 2 : c;
        ko[k] = q[k] >= 0 ? q[k] & 3 : 0;
    }
    int po[3];
    int32_t kk[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        po[k] = a.plotter[o[k]];
        kk[k] = a.kidx[o[k]];
    }
    // the neighbour's cell-edge on half s: the same end of the shared edge is
    // half s across it when the two run in opposite senses (types of
    // different signs), half 1 - s when they run in the same
    int32_t ev[3][2];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int so = ((pc > 0) == (po[k] > 0)) ? 1 - s : s;
            ev[k][s] = a.eslot[6 * o[k] + 2 * ko[k] + so];
        }
    asm volatile("" ::"v"(ev[0][0]), "v"(ev[0][1]), "v"(ev[1][0]), "v"(ev[1][1]), "v"(ev[2][0]),
                 "v"(ev[2][1]));
    int32_t y[3][2];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int s = 0; s < 2; ++s)
            y[k][s] = q[k] < 0 ? -1 : (kk[k] >= 0 ? (kk[k] << 2) | ko[k] : ev[k][s]);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (x[k][0] < 0) continue;
        for (int s = 0; s < 2; ++s) {
            const int32_t xs = x[k][s];
            if (s == 1 && xs == x[k][0]) break;
            const int32_t ys = y[k][s];
            if (ys >= 0) a.nbr_out[3 * (int64_t)(ys >> 2) + (ys & 3)] = xs;
            a.nbr_out[3 * (int64_t)(xs >> 2) + (xs & 3)] = ys;
        }
    }
}

// Transpose of the state transfer of a remeshing (a linear map): a split
// copied the parent to its children, so the parent's adjoint is the sum of
// the children's; a merge averaged a pair, so each of the pair gets half of
// the parent's.  b follows the same map, except that split cells which
// sampled the raster do not depend on the parent's b.  gb_old accumulates.
__global__ __launch_bounds__(AT) void transfer_adjoint_kernel(
    int64_t n_old, const int8_t *code, const int64_t *base, int b_sampled, const double *le,
    const double *lu, const double *lv, const double *gb, double *le_o, double *lu_o,
    double *lv_o, double *gb_o) {
    const int64_t c = blockIdx.x * (int64_t)AT + threadIdx.x;
    if (c >= n_old) return;
    const int8_t k = code[c];
    const int64_t b0 = base[c] & 0xffffffffLL;
    if (k == CODE_KEEP) {
        le_o[c] = le[b0];
        lu_o[c] = lu[b0];
        lv_o[c] = lv[b0];
        gb_o[c] += gb[b0];
    } else if (k == CODE_MERGE_FIRST || k == CODE_MERGE_SECOND) {
        const int64_t par = k == CODE_MERGE_FIRST ? b0 : b0 - 1;
        le_o[c] = 0.5 * le[par];
        lu_o[c] = 0.5 * lu[par];
        lv_o[c] = 0.5 * lv[par];
        gb_o[c] += 0.5 * gb[par];
    } else {
        const int pat = k - CODE_SPLIT;
        const int cells = 2 + (pat & 1) + (pat >> 1);
        double se = 0.0, su = 0.0, sv = 0.0, sb = 0.0;
        for (int j = 0; j < cells; ++j) {
            se += le[b0 + j];
            su += lu[b0 + j];
            sv += lv[b0 + j];
            sb += gb[b0 + j];
        }
        le_o[c] = se;
        lu_o[c] = su;
        lv_o[c] = sv;
        if (!b_sampled) gb_o[c] += sb;
    }
}

// code 0 and base = c: the record of a remeshing that changed nothing
__global__ __launch_bounds__(AT) void identity_record_kernel(int64_t n, int8_t *code,
                                                             int64_t *base) {
    const int64_t c = blockIdx.x * (int64_t)AT + threadIdx.x;
    if (c >= n) return;
    code[c] = CODE_KEEP;
    base[c] = c;
}

inline unsigned blocks_for(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

hipError_t scan_i64(const int64_t *in, int64_t *out, int64_t n, int64_t *bsum, hipStream_t s) {
    const int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nb > 0) {
        scan_reduce_kernel<<<(unsigned)nb, SCAN_T, 0, s>>>(in, n, bsum);
        scan_top_kernel<<<1, TOP_T, 0, s>>>(bsum, nb);
        scan_apply_kernel<<<(unsigned)nb, SCAN_T, 0, s>>>(in, out, n, bsum);
    } else {
        scan_top_kernel<<<1, TOP_T, 0, s>>>(bsum, 0);
    }
    return hipGetLastError();
}

// Refinement flags from the largest jump of the surface elevation across an
// edge to a wet neighbour (a wet cell only; h = eta - b > drytol).
__global__ __launch_bounds__(AT) void eta_jump_kernel(int64_t n, const double *eta,
                                                      const double *b, const int32_t *nbr,
                                                      const int8_t *depth, double drytol,
                                                      double refine_tol, double coarsen_tol,
                                                      int dmin, int dmax, int8_t *flags) {
    const int64_t c = blockIdx.x * (int64_t)AT + threadIdx.x;
    if (c >= n) return;
    const double e = eta[c];
    double j = 0.0;
    if (e - b[c] > drytol) {
        for (int k = 0; k < 3; ++k) {
            const int32_t q = nbr[3 * c + k];
            if (q < 0) continue;
            const int64_t o = q >> 2;
            const double eo = eta[o];
            if (eo - b[o] > drytol) j = fmax(j, fabs(e - eo));
        }
    }
    const int d = depth[c];
    int8_t f = 0;
    if (d < dmax && j > refine_tol) f = 1;
    else if (d > dmin && j < coarsen_tol) f = -1;
    flags[c] = f;
}

// res[0] = the covered part of the two root trees.
__global__ void copy_total_kernel(int64_t *dst, const int64_t *bsum, int64_t tiles) {
    if (threadIdx.x == 0) *dst = bsum[tiles];
}
}  // namespace

int64_t adapt_scan_tiles(int64_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }

// Waits until the device has stored seq at host[7]: polls the word for a
// while (a remeshing's first phase takes well under a millisecond), then
// blocks on the stream, which also reports a failed kernel.
static hipError_t wait_results(const int64_t *host, int64_t seq, hipStream_t s) {
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned i = 1;; ++i) {
        if (__atomic_load_n(&host[7], __ATOMIC_ACQUIRE) == seq) return hipSuccess;
        if ((i & 255u) == 0u &&
            std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(20))
            break;
        __builtin_ia32_pause();
    }
    const hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    return __atomic_load_n(&host[7], __ATOMIC_ACQUIRE) == seq ? hipSuccess : hipErrorUnknown;
}

hipError_t adapt_plan(const AdaptArgs &a, hipStream_t s, int first_batch, bool offsets_valid,
                      bool prefix_ahead, int max_sweeps, uint32_t epoch, int64_t *host,
                      int64_t *host_dev, int64_t *seq, AdaptPlan *plan) {
    const unsigned nb = blocks_for(a.n, SWEEP_TILE);  // conformity sweep tiles
    const int64_t tiles = adapt_scan_tiles(a.n);
    int64_t *res = a.bsum + tiles + 1;
    hipError_t e;
    if (!offsets_valid) {  // else the offsets of the unchanged mesh are in place
        tree_size_kernel<<<blocks_for(a.n, AT), AT, 0, s>>>(a.n, a.depth, a.off);
        if ((e = scan_i64(a.off, a.off, a.n, a.bsum, s)) != hipSuccess) return e;
        copy_total_kernel<<<1, 64, 0, s>>>(res, a.bsum, tiles);
    }
    // Sweeps go in batches, each followed by the count pass and one host
    // read: the count pass also tells whether the words are at the fixpoint,
    // so no sweep is spent on confirming it.
    plan->sweeps = -1;
    const uint32_t tag = sweep_number(epoch, 0);
    if (max_sweeps > 0xffff) max_sweeps = 0xffff;
    int batch = first_batch < 1 ? 1 : first_batch, done = 0;
    while (done < max_sweeps) {
        for (int j = 1; j <= batch; ++j)
            conf_sweep_kernel<<<nb, AT, 0, s>>>(a.n, a.depth, a.nbr, a.slots, a.stamp,
                                                a.status + 1, sweep_number(epoch, (uint32_t)(done + j)),
                                                a.flags, a.dmin, a.dmax);
        done += batch;
        if (tiles > 0)
            count_reduce_kernel<<<(unsigned)tiles, SCAN_T, 0, s>>>(
                a.n, a.depth, a.nbr, a.slots, a.off, a.code, a.bsum, a.status);
        plan_top_kernel<<<1, TOP_T, 0, s>>>(a.bsum, tiles, a.status, res, host_dev, ++*seq);
        // the prefix pass needs only the count pass: behind it, it runs
        // while the host reads the results (a batch short of the fixpoint
        // is counted again, and so is its prefix)
        if (prefix_ahead && tiles > 0)
            scan_codes_kernel<<<(unsigned)tiles, SCAN_T, 0, s>>>(a.n, a.code, a.bsum, a.base,
                                                                 a.kidx, a.list);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = wait_results(host, *seq, s)) != hipSuccess) return e;
        if (!(host[2] & 4)) {
            plan->sweeps = done;
            break;
        }
        batch = 2;
    }
    if (plan->sweeps < 0) return hipSuccess;
    plan->prefix_done = prefix_ahead;
    // a last change of an earlier call (another tag) means none in this one
    plan->last_change = ((uint32_t)host[3] >> 16) == (tag >> 16) ? (int)(host[3] & 0xffff) : 0;
    plan->tree_total = offsets_valid ? (int64_t(1) << 61) : host[0];
    plan->n_new = host[1];
    plan->status = (uint32_t)host[2] & 3u;
    plan->n_changing = host[5];
    return hipSuccess;
}

hipError_t flag_eta_jumps(int64_t n, const double *eta, const double *b, const int32_t *nbr,
                          const int8_t *depth, double drytol, double refine_tol,
                          double coarsen_tol, int dmin, int dmax, int8_t *flags, hipStream_t s) {
    eta_jump_kernel<<<blocks_for(n, AT), AT, 0, s>>>(n, eta, b, nbr, depth, drytol, refine_tol,
                                                    coarsen_tol, dmin, dmax, flags);
    return hipGetLastError();
}

hipError_t transfer_adjoint(int64_t n_old, const int8_t *code, const int64_t *base, int b_sampled,
                            const double *le, const double *lu, const double *lv, const double *gb,
                            double *le_o, double *lu_o, double *lv_o, double *gb_o,
                            hipStream_t s) {
    transfer_adjoint_kernel<<<blocks_for(n_old, AT), AT, 0, s>>>(n_old, code, base, b_sampled, le,
                                                                lu, lv, gb, le_o, lu_o, lv_o, gb_o);
    return hipGetLastError();
}

// The full base of a remeshing's record: a kept cell's is its new index.
__global__ __launch_bounds__(AT) void record_base_kernel(int64_t n, const int32_t *kidx,
                                                         const int64_t *base, int64_t *out) {
    const int64_t c = blockIdx.x * (int64_t)AT + threadIdx.x;
    if (c >= n) return;
    const int32_t k = kidx[c];
    out[c] = k >= 0 ? (int64_t)k : base[c];
}

hipError_t record_base(int64_t n, const int32_t *kidx, const int64_t *base, int64_t *out,
                       hipStream_t s) {
    record_base_kernel<<<blocks_for(n, AT), AT, 0, s>>>(n, kidx, base, out);
    return hipGetLastError();
}

hipError_t identity_record(int64_t n, int8_t *code, int64_t *base, hipStream_t s) {
    identity_record_kernel<<<blocks_for(n, AT), AT, 0, s>>>(n, code, base);
    return hipGetLastError();
}

hipError_t adapt_emit(const AdaptArgs &a, hipStream_t s, bool prefix_done) {
    const int64_t tiles = adapt_scan_tiles(a.n);
    if (tiles > 0 && !prefix_done)
        scan_codes_kernel<<<(unsigned)tiles, SCAN_T, 0, s>>>(a.n, a.code, a.bsum, a.base, a.kidx,
                                                             a.list);
    emit_keep_kernel<<<blocks_for(a.n, KEEP_T), KEEP_T, 0, s>>>(a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || a.n_changing == 0) return e;
    emit_change_kernel<<<blocks_for(a.n_changing, CHANGE_T), CHANGE_T, 0, s>>>(a);
    link_kernel<<<blocks_for(a.n_changing, CHANGE_T), CHANGE_T, 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace swe
