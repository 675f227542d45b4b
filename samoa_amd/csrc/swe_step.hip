// swe_step.hip -- the fused SWE Euler step for gfx950 (MI355X).
//
// One launch performs one traversal of SWE_euler_timestep.f90 over the whole
// mesh: the GeoClaw flux on every edge (skeleton_scalar_op :196-215 and
// bnd_skeleton_scalar_op :231-255), the cell update with the wet/dry limiter
// (cell_update_op :257-284, volume_op :307-355) and the CFL / refinement
// reductions (post_traversal_grid_op :111-131).
//
// Mapping.  A workgroup owns a chunk of CHUNK consecutive SFC cells.  It
// stages the chunk's state in LDS, builds a compact list of the edges it must
// solve (every edge to a lower-indexed in-chunk partner is solved once, edges
// to cells outside the chunk and domain-boundary edges by this chunk), solves
// them with one thread per list entry, scatters both sides' fluxes into LDS,
// then updates its cells and reduces the CFL candidate.  The Sierpinski order
// keeps almost every neighbour inside the chunk, so ~1.5 Riemann solves per
// cell are done instead of 3, and halo reads are L2 hits.
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

#include "swe_internal.hpp"
#include "swe_riemann.hpp"

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

template <int SOLVER>
__global__ __launch_bounds__(CHUNK) void euler_step_kernel(StepArgs a) {
    __shared__ double s_q[4][CHUNK];        // eta, hu, hv, b of the chunk
    __shared__ double s_f[4][3][CHUNK];     // flux (h, hu, hv, s) per (edge, cell)
    __shared__ int32_t s_nbr[3][CHUNK];     // packed neighbour (o << 2 | back edge)
    __shared__ int8_t s_plot[CHUNK];
    __shared__ uint16_t s_list[3 * CHUNK];  // (local cell << 2) | edge
    __shared__ int s_wsum[CHUNK / 64];
    __shared__ double s_wmin[CHUNK / 64];
    __shared__ int s_wref[CHUNK / 64];

    // XCD-aware chunk order: consecutive SFC chunks stay on one XCD's L2.
    const int64_t nchunks = (a.n + CHUNK - 1) / CHUNK;
    const int64_t q = nchunks / 8, r = nchunks % 8;
    const int64_t xcd = blockIdx.x % 8, i = blockIdx.x / 8;
    if (i >= q + (xcd < r ? 1 : 0)) return;
    const int64_t chunk = xcd * q + (xcd < r ? xcd : r) + i;

    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int64_t base = chunk * CHUNK;
    const int cnt = (int)((a.n - base) < CHUNK ? (a.n - base) : CHUNK);
    const bool active = t < cnt;
    const int64_t c = base + t;
    const double dt = a.device_dt ? a.slots->dt : a.dt;

    int32_t nb[3] = {-1, -1, -1};
    int8_t plot = 1;
    if (active) {
        s_q[0][t] = a.eta[c];
        s_q[1][t] = a.hu[c];
        s_q[2][t] = a.hv[c];
        s_q[3][t] = a.b[c];
        nb[0] = a.nbr[3 * c];
        nb[1] = a.nbr[3 * c + 1];
        nb[2] = a.nbr[3 * c + 2];
        plot = a.plotter[c];
    }
    s_nbr[0][t] = nb[0];
    s_nbr[1][t] = nb[1];
    s_nbr[2][t] = nb[2];
    s_plot[t] = plot;

    // Edges this thread's cell must solve: boundary, halo, or in-chunk with a
    // higher-indexed partner.
    int own = 0;
    if (active) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int32_t p = nb[k];
            const int64_t o = (int64_t)(p >> 2);
            const bool inside = p >= 0 && o >= base && o < base + cnt;
            if (!inside || o > c) own |= 1 << k;
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
    for (int w = 0; w < CHUNK / 64; ++w) {
        const int ws = s_wsum[w];
        pre += w < wave ? ws : 0;
        total += ws;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if (own & (1 << k)) s_list[pre++] = (uint16_t)((t << 2) | k);
    __syncthreads();

    const double g = a.g, drytol = a.drytol;
    for (int e = t; e < total; e += CHUNK) {
        const int ent = s_list[e];
        const int lc = ent >> 2, k = ent & 3;
        const int64_t cc = base + lc;
        const int32_t p = s_nbr[k][lc];
        const int pi = s_plot[lc] + 8;
        const double nx = c_normal[(pi * 3 + k) * 2], ny = c_normal[(pi * 3 + k) * 2 + 1];
        const double qe = s_q[0][lc], qu = s_q[1][lc], qv = s_q[2][lc], qb = s_q[3][lc];
        Flux fL, fR;
        if (p < 0) {  // domain boundary: ghost state (bnd_skeleton_scalar_op)
            double gu = qu, gv = qv;
            if (a.bc == 1) {
                const double dp = qu * nx + qv * ny;
                gu = qu - (2.0 * dp) * nx;
                gv = qv - (2.0 * dp) * ny;
            }
            edge_flux<SOLVER>(nx, ny, qe, qu, qv, qb, qe, gu, gv, qb, drytol, g, fL, fR);
            s_f[0][k][lc] = fL.h;
            s_f[1][k][lc] = fL.hu;
            s_f[2][k][lc] = fL.hv;
            s_f[3][k][lc] = fL.s;
            continue;
        }
        const int64_t o = (int64_t)(p >> 2);
        const int ko = p & 3;
        const bool inside = o >= base && o < base + cnt;
        double oe, ou, ov, ob;
        if (inside) {
            const int lo = (int)(o - base);
            oe = s_q[0][lo];
            ou = s_q[1][lo];
            ov = s_q[2][lo];
            ob = s_q[3][lo];
        } else {
            oe = a.eta[o];
            ou = a.hu[o];
            ov = a.hv[o];
            ob = a.b[o];
        }
        const bool c_left = a.forward ? (cc < o) : (cc > o);
        Flux fc, fo;
        if (c_left) {
            edge_flux<SOLVER>(nx, ny, qe, qu, qv, qb, oe, ou, ov, ob, drytol, g, fc, fo);
        } else {
            edge_flux<SOLVER>(-nx, -ny, oe, ou, ov, ob, qe, qu, qv, qb, drytol, g, fo, fc);
        }
        s_f[0][k][lc] = fc.h;
        s_f[1][k][lc] = fc.hu;
        s_f[2][k][lc] = fc.hv;
        s_f[3][k][lc] = fc.s;
        if (inside) {
            const int lo = (int)(o - base);
            s_f[0][ko][lo] = fo.h;
            s_f[1][ko][lo] = fo.hu;
            s_f[2][ko][lo] = fo.hv;
            s_f[3][ko][lo] = fo.s;
        }
    }
    __syncthreads();

    // cell update: volume_op + cell_update_op
    double cand = 1.7976931348623157e308;  // huge(1.0_SR), pre_traversal_op :139
    int refined = 0;
    if (active) {
        const int d = a.depth[c];
        const double S = a.scaling;
        const double volume = (S * S) * ldexp(1.0, -(d + 1));
        const double lleg = S * c_leg[d + 1], lhyp = S * c_leg[d];
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
        const double dq = fabs(dh);
        if (d < a.dmax && dq > a.refine_thr) {
            rf = 1;
            refined = 1;
        } else if (d > a.dmin && dq < a.refine_thr / 8.0) {
            rf = -1;
        }
        if (a.refine) a.refine[c] = rf;
        cand = volume / sv;
        const double scale = -dt / volume;
        dh = scale * dh;
        d1 = scale * d1;
        d2 = scale * d2;

        double qh = s_q[0][t], q1 = s_q[1][t], q2 = s_q[2][t];
        const double qb = s_q[3][t];
        if (qh < qb + drytol && dh > 0.0) {  // wetting, :271-275
            qh = qb + drytol;
            q1 = 0.0;
            q2 = 0.0;
        }
        qh = qh + dh;
        q1 = q1 + d1;
        q2 = q2 + d2;
        if (qh < qb + drytol) {  // drying, :280-283
            qh = dmin2(qb, 0.0);
            q1 = 0.0;
            q2 = 0.0;
        }
        a.eta_out[c] = qh;
        a.hu_out[c] = q1;
        a.hv_out[c] = q2;
    }

    // CFL minimum and refinement count (post_traversal_grid_op :117-118)
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
        for (int w = 1; w < CHUNK / 64; ++w) {
            m = dmin2(m, s_wmin[w]);
            rsum += s_wref[w];
        }
        Slots *sl = a.slots;
        // non-negative doubles order like their bit patterns
        atomicMin(&sl->dtmin, (unsigned long long)__double_as_longlong(m));
        if (rsum) atomicAdd(&sl->nref, (unsigned long long)rsum);
        __threadfence();  // publish this workgroup's contribution before counting it
        const unsigned int done = atomicAdd(&sl->counter, 1u);
        if (done == (unsigned int)(nchunks - 1)) {
            // last workgroup of the step: every contribution is visible
            __threadfence();
            const unsigned long long bits = atomicAdd(&sl->dtmin, 0ull);
            const unsigned long long nr = atomicAdd(&sl->nref, 0ull);
            const double mn = __longlong_as_double((long long)bits);
            sl->res_dtcfl = mn;
            sl->res_nref = nr;
            if (a.out_dtcfl) *a.out_dtcfl = mn;
            if (a.out_nref) *a.out_nref = (int64_t)nr;
            if (a.device_dt) {
                // post_traversal_grid_op :115-128 and the next pre_traversal_grid_op :93-99
                sl->time = sl->time + dt;
                double next = a.courant * mn;
                if (a.dt_clamp > 0.0) next = dmin2(a.dt_clamp, next);
                sl->dt = next;
            }
            atomicExch(&sl->dtmin, (unsigned long long)__double_as_longlong(1.7976931348623157e308));
            atomicExch(&sl->nref, 0ull);
            atomicExch(&sl->counter, 0u);
            __threadfence();
        }
    }
}

hipError_t upload_tables(const double *normal, const double *leg) {
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_normal), normal, sizeof(double) * 17 * 3 * 2);
    if (e != hipSuccess) return e;
    return hipMemcpyToSymbol(HIP_SYMBOL(c_leg), leg, sizeof(double) * 64);
}

hipError_t launch_step(const StepArgs &a, int solver, hipStream_t s) {
    const int64_t nchunks = (a.n + CHUNK - 1) / CHUNK;
    const int64_t blocks = ((nchunks + 7) / 8) * 8;
    if (solver == 1)
        hipLaunchKernelGGL(euler_step_kernel<1>, dim3((unsigned)blocks), dim3(CHUNK), 0, s, a);
    else
        hipLaunchKernelGGL(euler_step_kernel<0>, dim3((unsigned)blocks), dim3(CHUNK), 0, s, a);
    return hipGetLastError();
}

}  // namespace swe
