// swe_adjoint_step.hip -- discrete adjoint of one Euler traversal.
//
// For a traversal Q^{n+1} = F(Q^n, b; dt_n) (swe_step.hip), given the
// adjoint lam^{n+1} = dL/dQ^{n+1}, this kernel computes
//   lam^n  = (dF/dQ^n)^T lam^{n+1}        (eta, hu, hv)
//   gb    += (dF/db)^T   lam^{n+1}
// with dt_n frozen (taken from the forward pass) and the wet/dry limiter
// masks of the forward traversal frozen (cell_update_op,
// SWE_euler_timestep.f90:271-283: bit 0 = wetting fix taken, bit 1 = drying
// fix taken).
//
// Transpose update: the limiter passes the adjoint through, or stops it
// where a fix overwrote a value; the flux sum dQ = s * sum_k len_k f_k
// (s = -dt / vol) then seeds every edge flux of the cell with s * len_k * lam.
// Transpose flux: each edge of the chunk's compacted edge list (the same
// list the forward kernel builds) runs edge_flux_adj with the seeds of both
// cells, and the gradients with respect to both states are written to
// per-(cell, edge) LDS slots, summed per cell in a fixed order (so results do
// not depend on scheduling).  Edges to cells of other chunks are solved by
// both chunks, each keeping its own side.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "swe_adjoint.hpp"
#include "swe_internal.hpp"

namespace swe {

// Geometry tables of this translation unit (uploaded with the forward ones).
__constant__ double c_normal_adj[17 * 3 * 2];  // [plotter + 8][edge][x,y]
__constant__ double c_leg_adj[64];             // l(k) = sqrt(0.5)**k at k + 1

template <int SOLVER, int CH>
__global__ __launch_bounds__(CH) void adjoint_step_kernel(AdjArgs a) {
    __shared__ double s_q[4][CH];        // eta, hu, hv, b of the chunk at t_n
    __shared__ double s_seed[3][CH];     // s * lam^{n+1} after the limiter (per cell)
    __shared__ double s_gb[SOLVER != 1 ? CH : 1];  // each cell's b gradient so far
    __shared__ double s_g[4][3][CH];     // transpose-flux contributions per (edge, cell)
    __shared__ int32_t s_nbr[3][CH];
    __shared__ int8_t s_plot[CH];
    __shared__ int8_t s_dep[CH];
    __shared__ uint16_t s_list[3 * CH];
    __shared__ int s_wsum[CH / 64];

    __shared__ double s_normal[17 * 3 * 2];  // geometry tables in LDS (see swe_step.hip)
    __shared__ double s_leg[64];

    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const double g = a.g, inv16g = a.inv16g;
    for (int i = t; i < 17 * 3 * 2; i += CH) s_normal[i] = c_normal_adj[i];
    for (int i = t; i < 64; i += CH) s_leg[i] = c_leg_adj[i];
    __syncthreads();
    const int64_t nchunks = a.nchunks;
    const int64_t q = nchunks / 8, r = nchunks % 8;
    const int64_t nvirt = 8 * ((nchunks + 7) / 8);
    for (int64_t v = blockIdx.x; v < nvirt; v += gridDim.x) {
        const int64_t xcd = v % 8, i = v / 8;
        if (i >= q + (xcd < r ? 1 : 0)) continue;
        const int64_t chunk = xcd * q + (xcd < r ? xcd : r) + i;
        // raised wave priority for the chunk's loads and edge list, as in the
        // forward step kernel
        __builtin_amdgcn_s_setprio(2);
        const int s = a.chunk_scen ? a.chunk_scen[chunk] : 0;
        const Scen sc = a.scen[s];
        const int64_t lbase = (chunk - sc.chunk_off) * CH;
        const int cnt = (int)((sc.n - lbase) < CH ? (sc.n - lbase) : CH);
        const bool active = t < cnt;
        const int64_t lc_t = lbase + t;
        const int64_t gs = sc.cell_off, gt = sc.topo_off;
        const double dt = a.dt_dev ? a.dt_dev[s] : a.dt;
        const double drytol = sc.drytol, S = sc.scaling;

        // transpose of the cell update for this thread's cell
        int32_t nb[3] = {-1, -1, -1};
        int8_t plot = 1, dep = 0;
        double dir_e = 0.0, dir_u = 0.0, dir_v = 0.0, dir_b = 0.0;
        if (active) {
            // every load in one round: the state, topology, masks and
            // adjoint inputs first, the depth (behind a branch) last
            const int64_t x = gs + lc_t;
            const double qe = a.eta[x], qu = a.hu[x], qv = a.hv[x], qb = a.b[x];
            const int mk = a.masks[x];
            const double le = a.leta[x], lu = a.lhu[x], lv = a.lhv[x];
            // only this cell's thread adds to its b gradient: read now, with
            // the other loads, not after the edge loop (the f-wave kernel, at
            // 168 VGPRs and three waves per SIMD, reads it at the end: the
            // staged read costs it a register and a wave)
            if constexpr (SOLVER != 1) s_gb[t] = a.gb[x];
            nb[0] = a.nbr[3 * (gt + lc_t)];
            nb[1] = a.nbr[3 * (gt + lc_t) + 1];
            nb[2] = a.nbr[3 * (gt + lc_t) + 2];
            plot = a.plotter[gt + lc_t];
            dep = (int8_t)(sc.depth >= 0 ? sc.depth : a.depth[gt + lc_t]);
            s_q[0][t] = qe;
            s_q[1][t] = qu;
            s_q[2][t] = qv;
            s_q[3][t] = qb;
            double g2e = le, g2u = lu, g2v = lv;
            if (mk & 2) {  // drying fix: eta = min(b, 0), p = 0
                g2e = g2u = g2v = 0.0;
                if (qb < 0.0) dir_b += le;
            }
            if (mk & 1) {  // wetting fix: eta = b + drytol, p = 0 before the update
                dir_b += g2e;
            } else {
                dir_e = g2e;
                dir_u = g2u;
                dir_v = g2v;
            }
            const double vol = (S * S) * ldexp(1.0, -(dep + 1));
            const double scale = -dt / vol;
            s_seed[0][t] = scale * g2e;
            s_seed[1][t] = scale * g2u;
            s_seed[2][t] = scale * g2v;
        }
        s_nbr[0][t] = nb[0];
        s_nbr[1][t] = nb[1];
        s_nbr[2][t] = nb[2];
        s_plot[t] = plot;
        s_dep[t] = dep;
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int m = 0; m < 4; ++m) s_g[m][k][t] = 0.0;

        int own = 0;
        if (active) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const int32_t p = nb[k];
                const int64_t o = (int64_t)(p >> 2);
                const bool inside = p >= 0 && o >= lbase && o < lbase + cnt;
                if (!inside || o > lc_t) own |= 1 << k;
            }
        }
        const int nown = __popc(own);
        const uint64_t b0 = __ballot(nown & 1), b1 = __ballot(nown & 2);
        const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
        const int wpre = __popcll(b0 & lt) + 2 * __popcll(b1 & lt);
        if (lane == 0) s_wsum[wave] = __popcll(b0) + 2 * __popcll(b1);
        __syncthreads();
        int pre = wpre, total = 0;
#pragma unroll
        for (int w = 0; w < CH / 64; ++w) {
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
            const int ent = s_list[e];
            const int lc = ent >> 2, k = ent & 3;
            const int64_t cc = lbase + lc;
            const int32_t p = s_nbr[k][lc];
            const int pi = s_plot[lc] + 8;
            const double nx = s_normal[(pi * 3 + k) * 2], ny = s_normal[(pi * 3 + k) * 2 + 1];
            const double qe = s_q[0][lc], qu = s_q[1][lc], qv = s_q[2][lc], qb = s_q[3][lc];
            const int dc = s_dep[lc];
            const double len_c = S * (k == 1 ? s_leg[dc] : s_leg[dc + 1]);
            const double sc0 = s_seed[0][lc] * len_c, sc1 = s_seed[1][lc] * len_c,
                         sc2 = s_seed[2][lc] * len_c;
            const bool bnd = p < 0;
            const int64_t o = bnd ? cc : (int64_t)(p >> 2);
            const int ko = p & 3;
            const bool inside = !bnd && o >= lbase && o < lbase + cnt;
            double oe, ou, ov, ob, so0 = 0.0, so1 = 0.0, so2 = 0.0;
            if (bnd) {
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
                const int lo = (int)(o - lbase);
                oe = s_q[0][lo];
                ou = s_q[1][lo];
                ov = s_q[2][lo];
                ob = s_q[3][lo];
                const int dl = s_dep[lo];
                const double len_o = S * (ko == 1 ? s_leg[dl] : s_leg[dl + 1]);
                so0 = s_seed[0][lo] * len_o;
                so1 = s_seed[1][lo] * len_o;
                so2 = s_seed[2][lo] * len_o;
            } else {
                const int64_t x = gs + o;
                oe = a.eta[x];
                ou = a.hu[x];
                ov = a.hv[x];
                ob = a.b[x];
                const int mk = a.masks[x];
                // read with the rest, then selected (a read under the mask
                // test would wait for the mask first)
                const double l0 = a.leta[x], l1 = a.lhu[x], l2 = a.lhv[x];
                const int dl = sc.depth >= 0 ? sc.depth : a.depth[gt + o];
                // the neighbour's seed exactly as its own chunk forms it
                // ((s * lam) after the limiter, times the edge length), so
                // that results do not depend on the chunking
                const double vol = (S * S) * ldexp(1.0, -(dl + 1));
                const double scale = -dt / vol;
                const double len_o = S * (ko == 1 ? s_leg[dl] : s_leg[dl + 1]);
                const bool dry = (mk & 2) != 0;
                so0 = (scale * (dry ? 0.0 : l0)) * len_o;
                so1 = (scale * (dry ? 0.0 : l1)) * len_o;
                so2 = (scale * (dry ? 0.0 : l2)) * len_o;
            }
            const bool c_left = bnd || (a.forward ? (cc < o) : (cc > o));
            const double sx = c_left ? nx : -nx, sy = c_left ? ny : -ny;
            // seeds of the left and right sides, selected element-wise (a
            // pointer selection between two arrays would put them in scratch)
            const double gfl[3] = {c_left ? sc0 : so0, c_left ? sc1 : so1, c_left ? sc2 : so2};
            const double gfr[3] = {c_left ? so0 : sc0, c_left ? so1 : sc1, c_left ? so2 : sc2};
            double gQL[4] = {0, 0, 0, 0}, gQR[4] = {0, 0, 0, 0};
            edge_flux_adj<SOLVER>(sx, sy, c_left ? qe : oe, c_left ? qu : ou, c_left ? qv : ov,
                                  c_left ? qb : ob, c_left ? oe : qe, c_left ? ou : qu,
                                  c_left ? ov : qv, c_left ? ob : qb, drytol, g, inv16g,
                                  gfl, gfr, gQL, gQR);
            double gc[4], go[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                gc[m] = c_left ? gQL[m] : gQR[m];
                go[m] = c_left ? gQR[m] : gQL[m];
            }
            if (bnd) {  // the ghost is a function of this cell's state
                gc[0] += go[0];
                gc[3] += go[3];
                if (a.bc == 1) {
                    const double gdp = -2.0 * (go[1] * nx + go[2] * ny);
                    gc[1] += go[1] + gdp * nx;
                    gc[2] += go[2] + gdp * ny;
                } else {
                    gc[1] += go[1];
                    gc[2] += go[2];
                }
            }
#pragma unroll
            for (int m = 0; m < 4; ++m) s_g[m][k][lc] = gc[m];
            if (inside) {
                const int lo = (int)(o - lbase);
#pragma unroll
                for (int m = 0; m < 4; ++m) s_g[m][ko][lo] = go[m];
            }
        }
        __syncthreads();

        if (active) {
            const int64_t x = gs + lc_t;
            double le = dir_e, lu = dir_u, lv = dir_v, lb = dir_b;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                le += s_g[0][k][t];
                lu += s_g[1][k][t];
                lv += s_g[2][k][t];
                lb += s_g[3][k][t];
            }
            a.leta_out[x] = le;
            a.lhu_out[x] = lu;
            a.lhv_out[x] = lv;
            if constexpr (SOLVER != 1) a.gb[x] = s_gb[t] + lb;
            else a.gb[x] += lb;
        }
        __syncthreads();  // LDS is reused by the next chunk
    }
}

hipError_t upload_tables_adjoint(const double *normal, const double *leg) {
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_normal_adj), normal, sizeof(double) * 17 * 3 * 2);
    if (e != hipSuccess) return e;
    return hipMemcpyToSymbol(HIP_SYMBOL(c_leg_adj), leg, sizeof(double) * 64);
}

template <int SOLVER, int CH>
static void launch_adj_t(const AdjArgs &a, int64_t grid_blocks, hipStream_t s) {
    const int64_t nvirt = 8 * ((a.nchunks + 7) / 8);
    const int64_t blocks = grid_blocks > 0 && grid_blocks < nvirt ? grid_blocks : nvirt;
    hipLaunchKernelGGL((adjoint_step_kernel<SOLVER, CH>), dim3((unsigned)blocks), dim3(CH), 0, s,
                       a);
}

hipError_t launch_adjoint(const AdjArgs &a, int solver, int chunk, int64_t grid_blocks,
                          hipStream_t s) {
    if (grid_blocks < 0 || grid_blocks % 8) return hipErrorInvalidValue;
    if (solver != 0 && solver != 1) return hipErrorInvalidValue;  // the GeoClaw solvers only
    switch (chunk * 2 + (solver == 1 ? 1 : 0)) {
        case 128 * 2: launch_adj_t<0, 128>(a, grid_blocks, s); break;
        case 128 * 2 + 1: launch_adj_t<1, 128>(a, grid_blocks, s); break;
        case 256 * 2: launch_adj_t<0, 256>(a, grid_blocks, s); break;
        case 256 * 2 + 1: launch_adj_t<1, 256>(a, grid_blocks, s); break;
        case 512 * 2: launch_adj_t<0, 512>(a, grid_blocks, s); break;
        case 512 * 2 + 1: launch_adj_t<1, 512>(a, grid_blocks, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace swe
