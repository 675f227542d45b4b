// swe_step.hip -- the fused SWE Euler step for gfx950 (MI355X).
//
// One step kernel performs one traversal of SWE_euler_timestep.f90 over
// every mesh of a batch: the GeoClaw flux on every edge (skeleton_scalar_op
// :196-215 and bnd_skeleton_scalar_op :231-255), the cell update with the
// wet/dry limiter (cell_update_op :257-284, volume_op :307-355) and each
// chunk's CFL / refinement partials.  A small finalize kernel then reduces
// the partials per scenario and advances the dt chain
// (post_traversal_grid_op :111-131, pre_traversal_grid_op :89-109).  The step
// kernel uses no atomics: per-workgroup atomics on one per-scenario address
// serialise across the eight XCDs.
//
// Mapping.  A workgroup processes chunks of CH consecutive SFC cells of one
// scenario (one thread per cell).  It stages the chunk's state in LDS, builds
// a compact list of the edges it must solve (every edge to a higher-indexed
// in-chunk partner once, edges to cells outside the chunk and domain-boundary
// edges by this chunk), solves them with one thread per list entry, scatters
// both sides' fluxes into LDS, then updates its cells and reduces the CFL
// candidate.  The Sierpinski order keeps almost every neighbour inside the
// chunk, so ~1.5 Riemann solves per cell are done instead of 3, and halo
// reads are L2 hits.  Chunks are dealt to XCDs in contiguous SFC ranges so
// each XCD's L2 holds the halos of its own chunks; a grid smaller than the
// chunk count makes the launch persistent (grid-stride).
//
// Edge semantics (Tools_adaptive_traversal.f90:606-640).  The left state of
// an interior edge is the cell that touches it first in the current
// traversal direction: the lower SFC index when forward, the higher when
// backward; the edge normal is the left cell's outward normal.  The two cells
// of an edge see bit-identical normals of opposite sign, so each cell uses
// its own table entry and negates it when it is the right cell.  The update
// sums fluxes in the reference edge-index order of the traversal: edges
// 1,2,3 of the forward plotter type, or 3,2,1 when the cells are reversed
// (SFC_data_types.f90:473-486 negates the plotter type, swapping the legs).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "swe_internal.hpp"
#include "swe_riemann.hpp"
#include "swe_riemann_tol.hpp"

namespace swe {

__constant__ double c_normal[17 * 3 * 2];  // [plotter + 8][edge][x,y]
__constant__ double c_leg[64];             // l(k) = sqrt(0.5)**k at k + 1

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = dmin2(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ int wave_sum_i(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Four waves per SIMD (at most 128 VGPRs, a few spilled) outrun three
// unspilled ones: the LDS of four 256-cell workgroups fits a CU.
// JUMPS: the refinement flags follow the elevation-jump criterion
// (StepArgs::flag_jumps); a template parameter so that the plain kernel's
// code and registers do not pay for it.  RAW: one mesh (StepArgs::scen0)
// whose nbr holds plain neighbour cells (StepArgs::raw_nbr).  TOL: the
// toleranced arithmetic of swe_riemann_tol.hpp (StepArgs::tol; the GeoClaw
// solvers only): the chunk stages each cell's depth, world velocity,
// bathymetry and sqrt(g h) instead of its state and keeps its own state in
// registers.
template <int SOLVER, int CH, bool JUMPS, bool RAW = false, bool TOL = false>
__global__ __launch_bounds__(CH) __attribute__((amdgpu_waves_per_eu(4))) void euler_step_kernel(
    StepArgs a) {
    static_assert(!(RAW && TOL), "the toleranced kernel reads packed neighbours");
    static_assert(!TOL || SOLVER == 0 || SOLVER == 1, "toleranced GeoClaw solvers only");
    constexpr int NW = CH / 64;
    __shared__ double s_q[4][CH];        // eta, hu, hv, b of the chunk (TOL: h, wu, wv, b)
    __shared__ double s_x[TOL ? CH : 1];  // TOL: sqrt(g h) (0 when dry)
    __shared__ double s_f[4][3][CH];     // flux (h, hu, hv, s) per (edge, cell)
    __shared__ int32_t s_nbr[3][CH];     // packed neighbour (o << 2 | back edge)
    __shared__ int8_t s_plot[CH];
    __shared__ uint16_t s_list[3 * CH];  // (local cell << 2) | edge
    // JUMPS: elevation-jump bits per (edge, cell), from the edge loop, which
    // holds both states
    __shared__ uint8_t s_jb[JUMPS ? 3 : 1][JUMPS ? CH : 1];
    __shared__ int s_wsum[NW];
    __shared__ double s_wmin[NW];
    __shared__ int s_wref[NW];
    // geometry tables in LDS: the per-lane indexed reads would otherwise be
    // global loads that every edge waits for (the toleranced kernel reads
    // them from constant memory: its LDS holds four workgroups per CU only
    // without them)
    __shared__ double s_normal[TOL ? 1 : 17 * 3 * 2];
    __shared__ double s_leg[TOL ? 1 : 64];
    auto normal_at = [&](int i) -> double {
        if constexpr (TOL) return c_normal[i];
        else return s_normal[i];
    };
    auto leg_at = [&](int i) -> double {
        if constexpr (TOL) return c_leg[i];
        else return s_leg[i];
    };

    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const double g = a.g, inv16g = a.inv16g;
    if constexpr (!TOL) {
        for (int i = t; i < 17 * 3 * 2; i += CH) s_normal[i] = c_normal[i];
        for (int i = t; i < 64; i += CH) s_leg[i] = c_leg[i];
        __syncthreads();
    }

    // XCD-aware order: virtual block v runs on XCD v % 8 (the dispatcher deals
    // workgroups round robin and the grid is a multiple of 8); XCD x owns a
    // contiguous range of chunks.
    const int64_t nchunks = a.nchunks;
    const int64_t q = nchunks / 8, r = nchunks % 8;
    const int64_t nvirt = 8 * ((nchunks + 7) / 8);
    for (int64_t v = blockIdx.x; v < nvirt; v += gridDim.x) {
        const int64_t xcd = v % 8, i = v / 8;
        if (i >= q + (xcd < r ? 1 : 0)) continue;  // uniform across the workgroup
        const int64_t chunk = xcd * q + (xcd < r ? xcd : r) + i;
        // Raised wave priority from here to the edge loop: the chunk's
        // descriptor and staging loads, and the edge list, are a chain of
        // memory round trips and barriers; issued ahead of the other waves'
        // solver arithmetic, they overlap it (3-5 % off the step).
        __builtin_amdgcn_s_setprio(2);
        const int s = a.chunk_scen ? a.chunk_scen[chunk] : 0;
        const Scen sc = RAW ? a.scen0 : a.scen[s];
        // The state and topology arrays are addressed from per-scenario
        // (gs_*) and per-chunk (c_*) base pointers that are uniform across
        // the workgroup.  Cell indices are 64-bit in the exact kernel and
        // 32-bit in the toleranced one (a scenario holds at most 2^29 cells):
        // measured, each is the faster of the two for that kernel (by 3-4 %
        // and by about 1 %).
        using Index = std::conditional_t<TOL, int32_t, int64_t>;
        const Index lbase = (Index)((chunk - sc.chunk_off) * CH);  // first cell of the chunk
        const int cnt = (int)((sc.n - lbase) < CH ? (sc.n - lbase) : CH);
        const bool active = t < cnt;
        const int64_t gs = sc.cell_off, gt = sc.topo_off;  // state / topology offsets
        const double *const gs_eta = a.eta + gs, *const gs_hu = a.hu + gs, *const gs_hv = a.hv + gs;
        const double *const gs_b = a.b + gs;
        const double *const c_eta = gs_eta + lbase, *const c_hu = gs_hu + lbase;
        const double *const c_hv = gs_hv + lbase, *const c_b = gs_b + lbase;
        const int32_t *const c_nbr = a.nbr + 3 * (gt + lbase);
        const int8_t *const c_plot = a.plotter + gt + lbase;
        const int8_t *const c_depth = a.depth + gt + lbase;
        const int64_t c_out = gs + lbase;  // first output cell of the chunk
        const double dt = a.device_dt ? a.slots[s].dt : (a.dt_host ? a.dt_host[s] : a.dt);
        const double drytol = sc.drytol;

        int32_t nb[3] = {-1, -1, -1};
        int8_t plot = 1;
        double my_e = 0.0, my_u = 0.0, my_v = 0.0;  // TOL: this thread's cell state
        if (active) {
            if constexpr (TOL) {
                // the topology loads go out with the state's, before the
                // arithmetic that waits for the state
                my_e = c_eta[t];
                my_u = c_hu[t];
                my_v = c_hv[t];
                const double my_b = c_b[t];
                nb[0] = c_nbr[3 * t];
                nb[1] = c_nbr[3 * t + 1];
                nb[2] = c_nbr[3 * t + 2];
                plot = c_plot[t];
                const tol::CellQ cq = tol::cell_q(my_e, my_u, my_v, my_b, drytol, g);
                s_q[0][t] = cq.h;
                s_q[1][t] = cq.wu;
                s_q[2][t] = cq.wv;
                s_q[3][t] = cq.b;
                s_x[t] = cq.sq;
            } else {
                s_q[0][t] = c_eta[t];
                s_q[1][t] = c_hu[t];
                s_q[2][t] = c_hv[t];
                s_q[3][t] = c_b[t];
                nb[0] = c_nbr[3 * t];
                nb[1] = c_nbr[3 * t + 1];
                nb[2] = c_nbr[3 * t + 2];
                plot = c_plot[t];
            }
            if (RAW) {  // a neighbour outside the mesh reads as the boundary
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    if ((uint32_t)nb[k] >= (uint64_t)sc.n) nb[k] = -1;
            }
        }
        s_nbr[0][t] = nb[0];
        s_nbr[1][t] = nb[1];
        s_nbr[2][t] = nb[2];
        s_plot[t] = plot;

        // Edges this thread's cell must solve: boundary, halo, or in-chunk
        // with a higher-indexed partner.
        int own = 0;
        if (active) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const int32_t p = nb[k];
                if constexpr (TOL) {
                    const int lo = (p >> 2) - lbase;  // chunk-local neighbour
                    const bool inside = p >= 0 && (uint32_t)lo < (uint32_t)cnt;
                    if (!inside || lo > t) own |= 1 << k;
                } else {
                    const int64_t o = (int64_t)(RAW ? p : p >> 2);
                    const bool inside = p >= 0 && o >= lbase && o < lbase + cnt;
                    if (!inside || o > lbase + t) own |= 1 << k;
                }
            }
        }
        const int nown = __popc(own);
        // workgroup exclusive prefix of nown via ballots (nown in 0..3)
        const uint64_t b0 = __ballot(nown & 1), b1 = __ballot(nown & 2);
        const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
        const int wpre = __popcll(b0 & lt) + 2 * __popcll(b1 & lt);
        if (lane == 0) s_wsum[wave] = __popcll(b0) + 2 * __popcll(b1);
        __syncthreads();
        int pre = wpre, total = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            const int ws = s_wsum[w];
            pre += w < wave ? ws : 0;
            total += ws;
        }
#pragma unroll
        for (int k = 0; k < 3; ++k)
            if (own & (1 << k)) s_list[pre++] = (uint16_t)((t << 2) | k);
        __syncthreads();

        __builtin_amdgcn_s_setprio(0);
        for (int e = t; e < total; e += CH) {
            if constexpr (TOL) {
                const int ent = s_list[e];
                const int lc = ent >> 2, k = ent & 3;
                const int32_t p = s_nbr[k][lc];
                const int pi = s_plot[lc] + 8;
                const double nx = normal_at((pi * 3 + k) * 2), ny = normal_at((pi * 3 + k) * 2 + 1);
                const double hC = s_q[0][lc], uC = s_q[1][lc], vC = s_q[2][lc], bC = s_q[3][lc];
                const double sqC = s_x[lc];
                const bool bnd = p < 0;
                const int32_t o = RAW ? p : p >> 2;  // the neighbour (not read when bnd)
                const int lo = o - lbase;
                const bool inside = !bnd && (uint32_t)lo < (uint32_t)cnt;
                double hO, uO, vO, bO, sqO;
                if (bnd) {  // ghost: the cell itself, velocity mirrored at a wall
                    hO = hC;
                    bO = bC;
                    sqO = sqC;
                    uO = uC;
                    vO = vC;
                    if (a.bc == 1) {
                        const double dp = uC * nx + vC * ny;
                        uO = uC - (2.0 * dp) * nx;
                        vO = vC - (2.0 * dp) * ny;
                    }
                } else if (inside) {
                    hO = s_q[0][lo];
                    uO = s_q[1][lo];
                    vO = s_q[2][lo];
                    bO = s_q[3][lo];
                    sqO = s_x[lo];
                } else {
                    const tol::CellQ cq =
                        tol::cell_q(gs_eta[o], gs_hu[o], gs_hv[o], gs_b[o], drytol, g);
                    hO = cq.h;
                    uO = cq.wu;
                    vO = cq.wv;
                    bO = cq.b;
                    sqO = cq.sq;
                }
                // the left state as in the exact path: the operands of the
                // well-balancing sums (b + h of the right cell first) round
                // with the side they are on, and in deep water those
                // roundings are the largest part of a momentum update
                const bool c_left = bnd || (a.forward ? (lc < lo) : (lc > lo));
                const double sx = c_left ? nx : -nx, sy = c_left ? ny : -ny;
                Flux fL, fR;
                tol::edge_flux<SOLVER>(sx, sy, c_left ? hC : hO, c_left ? uC : uO,
                                       c_left ? vC : vO, c_left ? bC : bO, c_left ? sqC : sqO,
                                       c_left ? hO : hC, c_left ? uO : uC, c_left ? vO : vC,
                                       c_left ? bO : bC, c_left ? sqO : sqC, drytol, g, inv16g,
                                       fL, fR);
                const Flux fc = c_left ? fL : fR;
                const Flux fo = c_left ? fR : fL;
                s_f[0][k][lc] = fc.h;
                s_f[1][k][lc] = fc.hu;
                s_f[2][k][lc] = fc.hv;
                s_f[3][k][lc] = fc.s;
                uint8_t jb = 0;
                if constexpr (JUMPS) {
                    // as in the exact path below, with eta = h + b
                    if (!bnd && hC > drytol && hO > drytol) {
                        const double jd = fabs((hC + bC) - (hO + bO));
                        jb = (uint8_t)((jd > a.jump_refine ? 1 : 0) |
                                       (jd >= a.jump_coarsen ? 2 : 0));
                    }
                    s_jb[k][lc] = jb;
                }
                if (inside) {
                    const int ko = p & 3;
                    s_f[0][ko][lo] = fo.h;
                    s_f[1][ko][lo] = fo.hu;
                    s_f[2][ko][lo] = fo.hv;
                    s_f[3][ko][lo] = fo.s;
                    if constexpr (JUMPS) s_jb[ko][lo] = jb;
                }
                continue;
            }
            const int ent = s_list[e];
            const int lc = ent >> 2, k = ent & 3;
            const int64_t cc = lbase + lc;
            const int32_t p = s_nbr[k][lc];
            const int pi = s_plot[lc] + 8;
            const double nx = normal_at((pi * 3 + k) * 2), ny = normal_at((pi * 3 + k) * 2 + 1);
            const double qe = s_q[0][lc], qu = s_q[1][lc], qv = s_q[2][lc], qb = s_q[3][lc];
            const bool bnd = p < 0;
            const int64_t o = bnd ? cc : (int64_t)(RAW ? p : p >> 2);
            const bool inside = !bnd && o >= lbase && o < lbase + cnt;
            const int lo = (int)(o - lbase);  // chunk-local neighbour (used when inside)
            // the neighbour's edge back to this cell: packed with it, or for
            // plain neighbours found among the in-chunk neighbour's own
            // (two cells of a conforming mesh share at most one edge)
            int ko = p & 3;
            if (RAW && inside) ko = s_nbr[0][lo] == (int32_t)cc ? 0 : (s_nbr[1][lo] == (int32_t)cc ? 1 : 2);
            double oe, ou, ov, ob;
            if (bnd) {  // ghost state of bnd_skeleton_scalar_op (outflow copy or mirror)
                oe = qe;
                ob = qb;
                ou = qu;
                ov = qv;
                if (a.bc == 1) {
                    const double dp = qu * nx + qv * ny;
                    ou = qu - (2.0 * dp) * nx;
                    ov = qv - (2.0 * dp) * ny;
                }
            } else if (inside) {
                oe = s_q[0][lo];
                ou = s_q[1][lo];
                ov = s_q[2][lo];
                ob = s_q[3][lo];
            } else {
                oe = gs_eta[o];
                ou = gs_hu[o];
                ov = gs_hv[o];
                ob = gs_b[o];
            }
            // one solver call per entry: the left state is this cell on the
            // boundary and when it comes first in the traversal direction
            const bool c_left = bnd || (a.forward ? (cc < o) : (cc > o));
            const double sx = c_left ? nx : -nx, sy = c_left ? ny : -ny;
            Flux fL, fR;
            edge_flux<SOLVER>(sx, sy, c_left ? qe : oe, c_left ? qu : ou, c_left ? qv : ov,
                              c_left ? qb : ob, c_left ? oe : qe, c_left ? ou : qu,
                              c_left ? ov : qv, c_left ? ob : qb, drytol, g, inv16g, fL, fR);
            const Flux fc = c_left ? fL : fR;
            s_f[0][k][lc] = fc.h;
            s_f[1][k][lc] = fc.hu;
            s_f[2][k][lc] = fc.hv;
            s_f[3][k][lc] = fc.s;
            uint8_t jb = 0;
            if constexpr (JUMPS) {
                // the elevation jump across an edge between two wet cells:
                // bit 0 if above the refinement tolerance, bit 1 if not below
                // the coarsening tolerance (a NaN jump sets neither); the
                // same for both cells, as |a - b| = |b - a|
                if (!bnd && qe - qb > drytol && oe - ob > drytol) {
                    const double jd = fabs(qe - oe);
                    jb = (uint8_t)((jd > a.jump_refine ? 1 : 0) | (jd >= a.jump_coarsen ? 2 : 0));
                }
                s_jb[k][lc] = jb;
            }
            if (inside) {
                const Flux fo = c_left ? fR : fL;
                s_f[0][ko][lo] = fo.h;
                s_f[1][ko][lo] = fo.hu;
                s_f[2][ko][lo] = fo.hv;
                s_f[3][ko][lo] = fo.s;
                if constexpr (JUMPS) s_jb[ko][lo] = jb;
            }
        }
        __syncthreads();

        // cell update: volume_op + cell_update_op
        double cand = 1.7976931348623157e308;  // huge(1.0_SR), pre_traversal_op :139
        int refined = 0;
        if (active) {
            const int d = sc.depth >= 0 ? sc.depth : (int)c_depth[t];
            const double S = sc.scaling;
            const double volume = (S * S) * ldexp(1.0, -(d + 1));
            const double lleg = S * leg_at(d + 1), lhyp = S * leg_at(d);
            const int k0 = a.forward ? 0 : 2, k2 = a.forward ? 2 : 0;
            double dh = 0.0, d1 = 0.0, d2 = 0.0, sv = 0.0;
            dh = dh + lleg * s_f[0][k0][t];
            d1 = d1 + lleg * s_f[1][k0][t];
            d2 = d2 + lleg * s_f[2][k0][t];
            sv = sv + lleg * s_f[3][k0][t];
            dh = dh + lhyp * s_f[0][1][t];
            d1 = d1 + lhyp * s_f[1][1][t];
            d2 = d2 + lhyp * s_f[2][1][t];
            sv = sv + lhyp * s_f[3][1][t];
            dh = dh + lleg * s_f[0][k2][t];
            d1 = d1 + lleg * s_f[1][k2][t];
            d2 = d2 + lleg * s_f[2][k2][t];
            sv = sv + lleg * s_f[3][k2][t];

            int8_t rf = 0;
            if constexpr (JUMPS) {
                // elevation jumps to the wet neighbours of a wet cell, in
                // the state the step reads, from the edge loop's bits: bit 0
                // if one is above the refinement tolerance, bit 1 if one is
                // not below the coarsening tolerance.  j = the largest jump
                // (0 without a wet neighbour): j > tol if some edge's is,
                // j < tol if every edge's is and tol > 0.
                const int jb = s_jb[0][t] | s_jb[1][t] | s_jb[2][t];
                if (d < sc.dmax && (jb & 1)) {
                    rf = 1;
                    refined = 1;
                } else if (d > sc.dmin && !(jb & 2) && a.jump_coarsen > 0.0) {
                    rf = -1;
                }
            } else {
                const double dq = fabs(dh);
                if (d < sc.dmax && dq > sc.refine_thr) {
                    rf = 1;
                    refined = 1;
                } else if (d > sc.dmin && dq < sc.refine_thr / 8.0) {
                    rf = -1;
                }
            }
            if (a.refine) a.refine[c_out + t] = rf;
            cand = volume / sv;
            const double scale = -dt / volume;
            dh = scale * dh;
            d1 = scale * d1;
            d2 = scale * d2;

            double qh = TOL ? my_e : s_q[0][t], q1 = TOL ? my_u : s_q[1][t];
            double q2 = TOL ? my_v : s_q[2][t];
            const double qb = s_q[3][t];
            int mk = 0;
            if (qh < qb + drytol && dh > 0.0) {  // wetting, :271-275
                qh = qb + drytol;
                q1 = 0.0;
                q2 = 0.0;
                mk = 1;
            }
            qh = qh + dh;
            q1 = q1 + d1;
            q2 = q2 + d2;
            if (qh < qb + drytol) {  // drying, :280-283
                qh = dmin2(qb, 0.0);
                q1 = 0.0;
                q2 = 0.0;
                mk |= 2;
            }
            if (a.masks) a.masks[c_out + t] = (int8_t)mk;
            a.eta_out[c_out + t] = qh;
            a.hu_out[c_out + t] = q1;
            a.hv_out[c_out + t] = q2;
        }

        // CFL minimum and refinement count of the chunk (post_traversal_grid_op
        // :117-118); the finalize kernel reduces the chunks of each scenario.
        const double wmin = wave_min(cand);
        const int wref = wave_sum_i(refined);
        if (lane == 0) {
            s_wmin[wave] = wmin;
            s_wref[wave] = wref;
        }
        __syncthreads();
        if (t == 0) {
            double m = s_wmin[0];
            int rsum = s_wref[0];
#pragma unroll
            for (int w = 1; w < NW; ++w) {
                m = dmin2(m, s_wmin[w]);
                rsum += s_wref[w];
            }
            a.part_min[chunk] = m;
            a.part_nref[chunk] = rsum;
        }
        __syncthreads();  // LDS is reused by the next chunk
    }
}

// One workgroup per scenario: reduces the chunk partials of the step that
// just ran, publishes the results and advances the device dt chain
// (post_traversal_grid_op :115-128 and the next pre_traversal_grid_op :93-99).
// 1024 threads with four independent loads in flight each: the partials of
// a 4M-cell mesh (16k chunks) take four rounds of memory latency.
constexpr int FIN_THREADS = 1024;

// One workgroup per scenario.  Each thread reads FIN_PER partials of each
// array at once (every load at an index in range, the ones past the end
// unused), and the slot's time and dt are read at the start, so that the
// chain of memory round trips is short for meshes up to FIN_PER * FIN_THREADS
// chunks.
constexpr int FIN_PER = 16;
__global__ __launch_bounds__(FIN_THREADS) void finalize_kernel(StepArgs a) {
    __shared__ double s_min[FIN_THREADS / 64];
    __shared__ long long s_ref[FIN_THREADS / 64];
    const int s = blockIdx.x;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const Scen sc = a.scen ? a.scen[s] : a.scen0;
    Slots *sl = &a.slots[s];
    const double time0 = sl->time, dt0 = sl->dt;  // every thread: no branch to wait at
    const double huge = 1.7976931348623157e308;  // huge(1.0_SR)
    double m = huge;
    long long r = 0;
    const double *pm = a.part_min + sc.chunk_off;
    const int32_t *pr = a.part_nref + sc.chunk_off;
    const int64_t nc = sc.nchunks;
    for (int64_t i0 = t; i0 < nc; i0 += (int64_t)FIN_PER * FIN_THREADS) {
        double v[FIN_PER];
        int32_t w[FIN_PER];
#pragma unroll
        for (int j = 0; j < FIN_PER; ++j) {
            const int64_t i = i0 + (int64_t)j * FIN_THREADS;
            const int64_t ii = i < nc ? i : i0;
            v[j] = pm[ii];
            w[j] = pr[ii];
        }
#pragma unroll
        for (int j = 0; j < FIN_PER; ++j) {
            const bool in = i0 + (int64_t)j * FIN_THREADS < nc;
            m = dmin2(m, in ? v[j] : huge);
            r += in ? w[j] : 0;
        }
    }
    m = wave_min(m);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) r += __shfl_xor(r, off, 64);
    if (lane == 0) {
        s_min[wave] = m;
        s_ref[wave] = r;
    }
    __syncthreads();
    if (t == 0) {
        m = s_min[0];
        r = s_ref[0];
#pragma unroll
        for (int w = 1; w < FIN_THREADS / 64; ++w) {
            m = dmin2(m, s_min[w]);
            r += s_ref[w];
        }
        sl->res_dtcfl = m;
        sl->res_nref = r;
        if (a.out_dtcfl) a.out_dtcfl[s] = m;
        if (a.out_nref) a.out_nref[s] = (int64_t)r;
        if (a.device_dt) {
            sl->time = time0 + dt0;
            double next = sc.courant * m;
            if (sc.dt_clamp > 0.0) next = dmin2(sc.dt_clamp, next);
            sl->dt = next;
        }
    }
}

hipError_t upload_tables(const double *normal, const double *leg) {
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_normal), normal, sizeof(double) * 17 * 3 * 2);
    if (e != hipSuccess) return e;
    return hipMemcpyToSymbol(HIP_SYMBOL(c_leg), leg, sizeof(double) * 64);
}

template <int SOLVER, int CH>
static void launch_t(const StepArgs &a, int64_t grid_blocks, hipStream_t s) {
    const int64_t nvirt = 8 * ((a.nchunks + 7) / 8);
    const int64_t blocks = grid_blocks > 0 && grid_blocks < nvirt ? grid_blocks : nvirt;
    constexpr bool GEOCLAW = SOLVER == 0 || SOLVER == 1;  // the toleranced solvers
    if (GEOCLAW && a.tol) {
        if (a.flag_jumps)
            hipLaunchKernelGGL((euler_step_kernel<SOLVER, CH, true, false, GEOCLAW>),
                               dim3((unsigned)blocks), dim3(CH), 0, s, a);
        else
            hipLaunchKernelGGL((euler_step_kernel<SOLVER, CH, false, false, GEOCLAW>),
                               dim3((unsigned)blocks), dim3(CH), 0, s, a);
    } else if (a.flag_jumps)
        hipLaunchKernelGGL((euler_step_kernel<SOLVER, CH, true>), dim3((unsigned)blocks), dim3(CH),
                           0, s, a);
    else
        hipLaunchKernelGGL((euler_step_kernel<SOLVER, CH, false>), dim3((unsigned)blocks),
                           dim3(CH), 0, s, a);
    if (!a.skip_finalize)
        hipLaunchKernelGGL(finalize_kernel, dim3((unsigned)a.n_scen), dim3(FIN_THREADS), 0, s, a);
}

template <int SOLVER>
static hipError_t launch_chunk(const StepArgs &a, int chunk, int64_t grid_blocks, hipStream_t s) {
    switch (chunk) {
        case 128: launch_t<SOLVER, 128>(a, grid_blocks, s); break;
        case 256: launch_t<SOLVER, 256>(a, grid_blocks, s); break;
        case 512: launch_t<SOLVER, 512>(a, grid_blocks, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int SOLVER>
static void launch_raw_t(const StepArgs &a, hipStream_t s) {
    constexpr int CH = DEFAULT_CHUNK;
    const int64_t nvirt = 8 * ((a.nchunks + 7) / 8);
    hipLaunchKernelGGL((euler_step_kernel<SOLVER, CH, false, true>), dim3((unsigned)nvirt),
                       dim3(CH), 0, s, a);
    hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(FIN_THREADS), 0, s, a);
}

hipError_t launch_step_raw(const StepArgs &a, int solver, hipStream_t s) {
    if (a.scen || a.chunk_scen || a.n_scen != 1 || !a.raw_nbr || a.flag_jumps ||
        a.nchunks != (a.scen0.n + DEFAULT_CHUNK - 1) / DEFAULT_CHUNK)
        return hipErrorInvalidValue;
    switch (solver) {
        case 0: launch_raw_t<0>(a, s); break;
        case 1: launch_raw_t<1>(a, s); break;
        case SOLVER_LF: launch_raw_t<SOLVER_LF>(a, s); break;
        case SOLVER_LLF: launch_raw_t<SOLVER_LLF>(a, s); break;
        case SOLVER_LF_BATH: launch_raw_t<SOLVER_LF_BATH>(a, s); break;
        case SOLVER_LLF_BATH: launch_raw_t<SOLVER_LLF_BATH>(a, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_step(const StepArgs &a, int solver, int chunk, int64_t grid_blocks,
                       hipStream_t s) {
    if (a.raw_nbr || !a.scen) return hipErrorInvalidValue;
    if (grid_blocks < 0 || grid_blocks % 8) return hipErrorInvalidValue;
    switch (solver) {
        case 0: return launch_chunk<0>(a, chunk, grid_blocks, s);
        case 1: return launch_chunk<1>(a, chunk, grid_blocks, s);
        case SOLVER_LF: return launch_chunk<SOLVER_LF>(a, chunk, grid_blocks, s);
        case SOLVER_LLF: return launch_chunk<SOLVER_LLF>(a, chunk, grid_blocks, s);
        case SOLVER_LF_BATH: return launch_chunk<SOLVER_LF_BATH>(a, chunk, grid_blocks, s);
        case SOLVER_LLF_BATH: return launch_chunk<SOLVER_LLF_BATH>(a, chunk, grid_blocks, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace swe
