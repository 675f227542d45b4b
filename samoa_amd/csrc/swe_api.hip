// swe_api.hip -- host side of the C ABI declared in include/swe_mi355x.h.
//
// A `Batch` owns the device buffers of one or more frozen meshes (SoA state
// double-buffered and concatenated, topology shared between scenarios that
// pass the same mesh arrays), the per-scenario table the step kernel reads,
// and the reference's dt bookkeeping either on the host (swe_step,
// swe_batch_step) or on the device inside a captured HIP graph (swe_run,
// swe_batch_run).  A single-mesh handle is a batch of one.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "swe_internal.hpp"
#include "swe_sweep_stamp.hpp"
#include "swe_mi355x.h"
#include "swe_riemann.hpp"

namespace {

thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define HIP_TRY(expr)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(SWE_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));     \
    } while (0)

#define SWE_TRY(expr)                \
    do {                             \
        int rc_ = (expr);            \
        if (rc_ != SWE_OK) return rc_; \
    } while (0)


// Geometry tables, computed exactly as the reference initialises them.
// SFC_data_types.f90:627-686: Jacobian of each plotter type rounded to whole
// numbers, outward normals n_i = normalize(e_i J^-1), e = (-1,0),(1,1),(0,-1).
void normal_table(double out[17 * 3 * 2]) {
    const double PI = 4.0 * std::atan(1.0);
    std::memset(out, 0, sizeof(double) * 17 * 3 * 2);
    for (int p = -8; p <= 8; ++p) {
        if (p == 0) continue;
        const double a = PI / 4.0 * (double)std::abs(p);
        double J[2][2];
        if (p > 0) {
            J[0][0] = std::cos(a);
            J[1][0] = std::sin(a);
            J[0][1] = -std::sin(a);
            J[1][1] = std::cos(a);
        } else {
            J[0][0] = -std::sin(a);
            J[1][0] = std::cos(a);
            J[0][1] = std::cos(a);
            J[1][1] = std::sin(a);
        }
        for (auto &row : J)
            for (double &v : row) v = std::trunc(1.5 * v);
        const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        const double inv = 1.0 / det;
        const double Ji00 = inv * J[1][1], Ji10 = inv * (-J[1][0]);
        const double Ji01 = inv * (-J[0][1]), Ji11 = inv * J[0][0];
        const double e[3][2] = {{-1.0, 0.0}, {1.0, 1.0}, {0.0, -1.0}};
        for (int k = 0; k < 3; ++k) {
            const double x = e[k][0] * Ji00 + e[k][1] * Ji10;
            const double y = e[k][0] * Ji01 + e[k][1] * Ji11;
            const double f = 1.0 / std::sqrt(x * x + y * y);
            out[((p + 8) * 3 + k) * 2] = f * x;
            out[((p + 8) * 3 + k) * 2 + 1] = f * y;
        }
    }
}

// SFC_data_types.f90:557: sqrt(0.5)**k folded by binary exponentiation.
void leg_table(double out[64]) {
    const double s = std::sqrt(0.5);
    for (int k = -1; k <= 62; ++k) {
        int e = k < 0 ? -k : k;
        double r = 1.0, base = s;
        while (e) {
            if (e & 1) r = r * base;
            base = base * base;
            e >>= 1;
        }
        out[k + 1] = k < 0 ? 1.0 / r : r;
    }
}

// Devices the library serves (ordinals 0 .. MAX_DEVICES - 1); per-device
// set-up is recorded in fixed tables, so that it allocates nothing.
constexpr int MAX_DEVICES = 64;
std::mutex g_table_mutex;
bool g_tables_on_device[MAX_DEVICES] = {};

int ensure_tables(int device) {
    if (device < 0 || device >= MAX_DEVICES)
        return fail(SWE_ERR_ARG, "device %d beyond the %d the library serves", device, MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_table_mutex);
    if (g_tables_on_device[device]) return SWE_OK;
    double normal[17 * 3 * 2], leg[64];
    normal_table(normal);
    leg_table(leg);
    HIP_TRY(swe::upload_tables(normal, leg));
    HIP_TRY(swe::upload_tables_adjoint(normal, leg));
    g_tables_on_device[device] = true;
    return SWE_OK;
}

// The captured pair of steps bakes in which buffer holds the state, the
// starting direction and the launch shape; a pair returns the state to the
// buffer it started from, so the graph stays valid across run calls.  A
// handle keeps one per (state buffer, starting direction): a run of an odd
// number of steps leaves the next run in the other buffer or direction, which
// would otherwise capture again.
struct GraphCache {
    hipGraphExec_t exec = nullptr;
    int cur = -1, forward = -1, chunk = 0, keep = -1, jumps = -1, tol = -1;
    double jump_refine = 0.0, jump_coarsen = 0.0;
    int64_t grid = -1;
};

// Validates a mesh and packs its neighbours as (o << 2) | back edge.
int pack_mesh(const swe_mesh_desc &d, const double *normal, std::vector<int32_t> &packed,
              int *uniform_depth) {
    const int64_t n = d.n_cells;
    packed.assign((size_t)(3 * n), -1);
    int ud = d.depth[0];
    // every cell's depth and plotter type first: the topology pass reads the
    // neighbours' plotter types to index the normal table
    for (int64_t c = 0; c < n; ++c) {
        const int dep = d.depth[c];
        const int p = d.plotter[c];
        if (dep < 0 || dep > 60) return fail(SWE_ERR_ARG, "depth[%lld] = %d", (long long)c, dep);
        if (dep != ud) ud = -1;
        if (p < -8 || p > 8 || p == 0)
            return fail(SWE_ERR_ARG, "plotter[%lld] = %d", (long long)c, p);
        for (int k = 0; k < 3; ++k) {
            const int32_t o = d.nbr[3 * c + k];
            if (o < -1 || o >= n || o == c)
                return fail(SWE_ERR_ARG, "nbr[%lld] = %d invalid", (long long)(3 * c + k), o);
        }
    }
    for (int64_t c = 0; c < n; ++c) {
        const int p = d.plotter[c];
        for (int k = 0; k < 3; ++k) {
            const int32_t o = d.nbr[3 * c + k];
            if (o < 0) continue;
            int ko = -1;
            for (int j = 0; j < 3; ++j)
                if (d.nbr[3 * (int64_t)o + j] == (int32_t)c) ko = j;
            if (ko < 0)
                return fail(SWE_ERR_ARG, "neighbour relation not symmetric at cell %lld edge %d",
                            (long long)c, k);
            const int po = d.plotter[o];
            const double *nc = &normal[((p + 8) * 3 + k) * 2];
            const double *no = &normal[((po + 8) * 3 + ko) * 2];
            if (!(nc[0] == -no[0] && nc[1] == -no[1]))
                return fail(SWE_ERR_ARG,
                            "cells %lld and %d disagree on the normal of their shared edge",
                            (long long)c, o);
            packed[3 * c + k] = (int32_t)((o << 2) | ko);
        }
        // Sierpinski order: consecutive cells share an edge.  The kernels rely
        // on it: a run of m consecutive cells then has at most m + 2 edges
        // leaving it, which bounds the LDS halo of a chunk.
        if (c + 1 < n && d.nbr[3 * c] != c + 1 && d.nbr[3 * c + 1] != c + 1 &&
            d.nbr[3 * c + 2] != c + 1)
            return fail(SWE_ERR_ARG,
                        "cells %lld and %lld are consecutive but share no edge (not in "
                        "Sierpinski order)", (long long)c, (long long)(c + 1));
    }
    *uniform_depth = ud;
    return SWE_OK;
}

}  // namespace

struct Batch {
    int device = 0;
    hipStream_t stream = nullptr;      // where work is launched (caller may replace it)
    hipStream_t own_stream = nullptr;  // created by the library; used for graph capture
    int solver = 0, bc = 0;
    int tol = 0;  // SWE_ARITH_TOLERANCED: the toleranced step kernels
    double g = 9.80665;
    int32_t nscen = 0;
    std::vector<swe::Scen> scen;
    std::vector<int> uniform_depth;
    int64_t ncells = 0, ntopo = 0, nchunks = 0;
    int chunk = swe::DEFAULT_CHUNK;
    int64_t grid = 0;
    swe::Scen *d_scen = nullptr;
    int32_t *d_chunk_scen = nullptr;
    swe::Slots *d_slots = nullptr;
    double *d_dt = nullptr;  // [nscen] host-chosen dt per scenario
    double *d_part_min = nullptr;     // [part_cap] CFL minimum per chunk
    int32_t *d_part_nref = nullptr;   // [part_cap] refinement count per chunk
    int64_t part_cap = 0;
    double *eta[2] = {nullptr, nullptr}, *hu[2] = {nullptr, nullptr}, *hv[2] = {nullptr, nullptr};
    double *b = nullptr;
    int32_t *nbr = nullptr;
    int8_t *plotter = nullptr, *depth = nullptr, *refine = nullptr;
    int cur = 0;
    std::vector<char> uploaded;
    GraphCache graph[2][2];  // [state buffer][starting direction]
    // remeshing (single mesh): the state and topology buffers hold `cap`
    // cells; a spare topology and bathymetry set receives a new mesh that
    // fits, the spare state buffer eta/hu/hv[1 - cur] its state
    int64_t cap = 0;
    double *b2 = nullptr;
    int32_t *nbr2 = nullptr;
    int8_t *plotter2 = nullptr, *depth2 = nullptr;
    int64_t cap2 = 0;
    uint32_t *ad_slots = nullptr, *ad_status = nullptr;
    int64_t *ad_off = nullptr, *ad_base = nullptr, *ad_bsum = nullptr;
    int64_t *ad_off2 = nullptr;  // the new mesh's tree offsets, swapped with ad_off
    int64_t ad_off_cap = 0, ad_off2_cap = 0;
    int8_t *ad_code = nullptr;
    int32_t *ad_kidx = nullptr;  // new index of each kept cell, -1 for a changing one
    int32_t *ad_list = nullptr;  // the changing cells
    int32_t *ad_eslot = nullptr;
    int64_t ad_cap = 0;
    bool ad_off_valid = false;  // ad_off holds the current mesh's tree offsets
    int64_t *ad_host = nullptr; // [8] pinned results of the remesher
    int64_t *ad_host_dev = nullptr;  // its device address
    int64_t ad_seq = 0;         // number of the last results
    int ad_last_sweeps = 2;     // sweeps that changed a word in the last remeshing
    bool ad_last_changed = false;  // the last remeshing changed the mesh
    uint32_t ad_epoch = 0;      // number of the current remeshing
    int64_t ad_record_n = -1;      // old cells of the last remeshing (its record in ad_code/ad_base)
    bool ad_record_sampled = false;  // its split cells sampled the raster
    double *raster = nullptr;  // bathymetry raster for split cells (swe_set_bathymetry_raster)
    int64_t rnx = 0, rny = 0;
    double rx0 = 0, ry0 = 0, rdx = 1, rdy = 1, roffset_x = 0, roffset_y = 0;
    int keep_flags = 0;       // every step writes its refinement flags to `refine`
    int jump_flags = 0;       // ... from elevation jumps (swe_keep_eta_jump_flags)
    double jump_refine = 0.0, jump_coarsen = 0.0;
    bool flags_valid = false; // `refine` holds the flags of the current state's last step
};

struct swe_handle {
    Batch B;
};
struct swe_batch {
    Batch B;
};

namespace {

void release(Batch &B) {
    // teardown: errors cannot be reported from here, the buffers go either way
    for (auto &row : B.graph)
        for (GraphCache &gc : row)
            if (gc.exec) (void)hipGraphExecDestroy(gc.exec);
    for (int i = 0; i < 2; ++i) {
        (void)hipFree(B.eta[i]);
        (void)hipFree(B.hu[i]);
        (void)hipFree(B.hv[i]);
    }
    (void)hipFree(B.b);
    (void)hipFree(B.nbr);
    (void)hipFree(B.plotter);
    (void)hipFree(B.depth);
    (void)hipFree(B.refine);
    (void)hipFree(B.d_scen);
    (void)hipFree(B.d_chunk_scen);
    (void)hipFree(B.d_slots);
    (void)hipFree(B.d_dt);
    (void)hipFree(B.d_part_min);
    (void)hipFree(B.d_part_nref);
    (void)hipFree(B.b2);
    (void)hipFree(B.nbr2);
    (void)hipFree(B.plotter2);
    (void)hipFree(B.depth2);
    (void)hipFree(B.ad_slots);
    (void)hipFree(B.ad_status);
    (void)hipFree(B.ad_off);
    (void)hipFree(B.ad_off2);
    (void)hipFree(B.ad_base);
    (void)hipFree(B.ad_bsum);
    (void)hipFree(B.ad_code);
    (void)hipFree(B.ad_kidx);
    (void)hipFree(B.ad_list);
    (void)hipFree(B.ad_eslot);
    (void)hipFree(B.raster);
    if (B.ad_host) (void)hipHostFree(B.ad_host);
    if (B.own_stream) (void)hipStreamDestroy(B.own_stream);
}

bool all_uploaded(const Batch &B) {
    return std::all_of(B.uploaded.begin(), B.uploaded.end(), [](char c) { return c != 0; });
}

// Chunk tables for the current chunk size: chunk offsets per scenario and the
// chunk -> scenario map (not needed for a single scenario).
// upload = false: the device copy of the scenario table is already
// current (the remesher's emit kernel wrote it).
int rebuild_chunks(Batch &B, bool upload = true) {
    int64_t off = 0;
    for (auto &s : B.scen) {
        s.chunk_off = off;
        s.nchunks = (s.n + B.chunk - 1) / B.chunk;
        off += s.nchunks;
    }
    B.nchunks = off;
    if (B.nchunks >= (int64_t(1) << 31))
        return fail(SWE_ERR_ARG, "too many chunks (%lld)", (long long)B.nchunks);
    if (B.nchunks > B.part_cap) {  // with room to grow: remeshing changes the count
        const int64_t cap = B.nchunks + B.nchunks / 4;
        if (B.stream) HIP_TRY(hipStreamSynchronize(B.stream));
        (void)hipFree(B.d_part_min);
        (void)hipFree(B.d_part_nref);
        B.d_part_min = nullptr;
        B.d_part_nref = nullptr;
        B.part_cap = 0;
        if (hipMalloc(&B.d_part_min, sizeof(double) * (size_t)cap) != hipSuccess ||
            hipMalloc(&B.d_part_nref, sizeof(int32_t) * (size_t)cap) != hipSuccess)
            return fail(SWE_ERR_NOMEM, "chunk partials allocation failed");
        B.part_cap = cap;
    }
    (void)hipFree(B.d_chunk_scen);
    B.d_chunk_scen = nullptr;
    if (B.nscen > 1) {
        std::vector<int32_t> map((size_t)B.nchunks);
        for (int32_t s = 0; s < B.nscen; ++s)
            std::fill(map.begin() + B.scen[s].chunk_off,
                      map.begin() + B.scen[s].chunk_off + B.scen[s].nchunks, s);
        if (hipMalloc(&B.d_chunk_scen, sizeof(int32_t) * map.size()) != hipSuccess)
            return fail(SWE_ERR_NOMEM, "chunk table allocation failed");
        HIP_TRY(hipMemcpy(B.d_chunk_scen, map.data(), sizeof(int32_t) * map.size(),
                          hipMemcpyHostToDevice));
    }
    if (upload)
        HIP_TRY(hipMemcpyAsync(B.d_scen, B.scen.data(), sizeof(swe::Scen) * B.scen.size(),
                               hipMemcpyHostToDevice, B.stream));
    for (auto &row : B.graph)
        for (GraphCache &gc : row)
            if (gc.exec) {
                (void)hipGraphExecDestroy(gc.exec);
                gc = GraphCache{};
            }
    return SWE_OK;
}

int batch_init(Batch &B, int32_t nscen, const swe_mesh_desc *descs) {
    if (nscen <= 0 || !descs) return fail(SWE_ERR_ARG, "no scenarios");
    B.nscen = nscen;
    B.solver = descs[0].solver;
    B.bc = descs[0].boundary;
    B.g = descs[0].g;
    double normal[17 * 3 * 2], leg[64];
    normal_table(normal);
    leg_table(leg);
    // unique topologies, keyed by the caller's mesh arrays
    std::map<std::tuple<const void *, const void *, const void *, int64_t>, int64_t> topo_off;
    std::vector<std::vector<int32_t>> packs;
    std::vector<const swe_mesh_desc *> topo_src;
    B.scen.resize((size_t)nscen);
    B.uniform_depth.resize((size_t)nscen);
    for (int32_t s = 0; s < nscen; ++s) {
        const swe_mesh_desc &d = descs[s];
        const int64_t n = d.n_cells;
        if (n <= 0 || n > (int64_t(1) << 29))
            return fail(SWE_ERR_ARG, "scenario %d: n_cells %lld out of range (1..2^29)", s,
                        (long long)n);
        if (!d.depth || !d.plotter || !d.nbr)
            return fail(SWE_ERR_ARG, "scenario %d: null mesh array", s);
        if (!(d.scaling > 0.0) || !(d.g >= swe::MIN_G && d.g <= swe::MAX_G) ||
            !(d.dry_tolerance >= swe::MIN_DRY_TOLERANCE))
            return fail(SWE_ERR_ARG,
                        "scenario %d: scaling must be > 0, g in [%g, %g], dry_tolerance >= %g", s,
                        swe::MIN_G, swe::MAX_G, swe::MIN_DRY_TOLERANCE);
        if (d.solver < SWE_SOLVER_AUG_RIEMANN || d.solver > SWE_SOLVER_LLF_BATH)
            return fail(SWE_ERR_ARG, "scenario %d: unknown solver %d", s, d.solver);
        if (d.boundary != SWE_BC_OUTFLOW && d.boundary != SWE_BC_WALL)
            return fail(SWE_ERR_ARG, "scenario %d: unknown boundary %d", s, d.boundary);
        if (d.solver != B.solver || d.boundary != B.bc || d.g != B.g)
            return fail(SWE_ERR_ARG, "scenario %d: solver, boundary and g must match scenario 0", s);
        if (d.max_depth < 0 || d.max_depth > 60)
            return fail(SWE_ERR_ARG, "scenario %d: max_depth %d out of range", s, d.max_depth);
        auto key = std::make_tuple((const void *)d.nbr, (const void *)d.plotter,
                                   (const void *)d.depth, n);
        auto it = topo_off.find(key);
        int ud;
        if (it == topo_off.end()) {
            packs.emplace_back();
            int rc = pack_mesh(d, normal, packs.back(), &ud);
            if (rc != SWE_OK) return fail(rc, "scenario %d: %s", s, g_last_error.c_str());
            topo_off[key] = B.ntopo;
            topo_src.push_back(&d);
            B.scen[s].topo_off = B.ntopo;
            B.ntopo += n;
        } else {
            B.scen[s].topo_off = it->second;
            ud = d.depth[0];
            for (int64_t c = 1; c < n && ud >= 0; ++c)
                if (d.depth[c] != ud) ud = -1;
        }
        swe::Scen &sc = B.scen[s];
        sc.cell_off = B.ncells;
        sc.n = n;
        sc.scaling = d.scaling;
        sc.drytol = d.dry_tolerance;
        sc.refine_thr = (5.0 * d.scaling) * leg[d.max_depth + 1];  // volume_op :339
        sc.courant = d.courant;
        sc.dt_clamp = -1.0;
        sc.depth = ud;
        sc.dmin = d.min_depth;
        sc.dmax = d.max_depth;
        B.ncells += n;
    }
    HIP_TRY(hipGetDevice(&B.device));
    SWE_TRY(ensure_tables(B.device));
    const size_t bytes = sizeof(double) * (size_t)B.ncells;
    bool ok = hipStreamCreate(&B.own_stream) == hipSuccess;
    B.stream = B.own_stream;
    for (int i = 0; i < 2 && ok; ++i)
        ok = hipMalloc(&B.eta[i], bytes) == hipSuccess && hipMalloc(&B.hu[i], bytes) == hipSuccess &&
             hipMalloc(&B.hv[i], bytes) == hipSuccess;
    ok = ok && hipMalloc(&B.b, bytes) == hipSuccess &&
         hipMalloc(&B.nbr, sizeof(int32_t) * 3 * (size_t)B.ntopo) == hipSuccess &&
         hipMalloc(&B.plotter, (size_t)B.ntopo) == hipSuccess &&
         hipMalloc(&B.depth, (size_t)B.ntopo) == hipSuccess &&
         hipMalloc(&B.refine, (size_t)B.ncells) == hipSuccess &&
         hipMalloc(&B.d_scen, sizeof(swe::Scen) * (size_t)nscen) == hipSuccess &&
         hipMalloc(&B.d_slots, sizeof(swe::Slots) * (size_t)nscen) == hipSuccess &&
         hipMalloc(&B.d_dt, sizeof(double) * (size_t)nscen) == hipSuccess;
    if (!ok) return fail(SWE_ERR_NOMEM, "device allocation of %lld cells failed", (long long)B.ncells);
    for (size_t i = 0; i < packs.size(); ++i) {
        const swe_mesh_desc &d = *topo_src[i];
        const int64_t off = topo_off[std::make_tuple((const void *)d.nbr, (const void *)d.plotter,
                                                     (const void *)d.depth, d.n_cells)];
        HIP_TRY(hipMemcpy(B.nbr + 3 * off, packs[i].data(), sizeof(int32_t) * packs[i].size(),
                          hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(B.plotter + off, d.plotter, (size_t)d.n_cells, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(B.depth + off, d.depth, (size_t)d.n_cells, hipMemcpyHostToDevice));
    }
    std::vector<swe::Slots> init((size_t)nscen);
    HIP_TRY(hipMemcpy(B.d_slots, init.data(), sizeof(swe::Slots) * init.size(),
                      hipMemcpyHostToDevice));
    B.uploaded.assign((size_t)nscen, 0);
    B.cap = B.ncells;
    return rebuild_chunks(B);
}

swe::StepArgs make_args(const Batch &B, int from, int to, int forward) {
    swe::StepArgs a{};
    a.eta = B.eta[from];
    a.hu = B.hu[from];
    a.hv = B.hv[from];
    a.b = B.b;
    a.eta_out = B.eta[to];
    a.hu_out = B.hu[to];
    a.hv_out = B.hv[to];
    a.nbr = B.nbr;
    a.plotter = B.plotter;
    a.depth = B.depth;
    a.scen = B.d_scen;
    a.chunk_scen = B.d_chunk_scen;
    a.slots = B.d_slots;
    a.part_min = B.d_part_min;
    a.part_nref = B.d_part_nref;
    a.n_scen = B.nscen;
    a.nchunks = B.nchunks;
    a.g = B.g;
    a.inv16g = 1.0 / (16.0 * B.g);
    a.bc = B.bc;
    a.forward = forward ? 1 : 0;
    a.flag_jumps = B.jump_flags;
    a.jump_refine = B.jump_refine;
    a.jump_coarsen = B.jump_coarsen;
    a.tol = B.tol;
    return a;
}

int batch_upload(Batch &B, int32_t s, const double *eta, const double *hu, const double *hv,
                 const double *b) {
    if (s < 0 || s >= B.nscen) return fail(SWE_ERR_ARG, "scenario %d out of range", s);
    if (!eta || !hu || !hv || !b) return fail(SWE_ERR_ARG, "upload: null argument");
    const size_t bytes = sizeof(double) * (size_t)B.scen[s].n;
    const int64_t o = B.scen[s].cell_off;
    HIP_TRY(hipMemcpyAsync(B.eta[B.cur] + o, eta, bytes, hipMemcpyHostToDevice, B.stream));
    HIP_TRY(hipMemcpyAsync(B.hu[B.cur] + o, hu, bytes, hipMemcpyHostToDevice, B.stream));
    HIP_TRY(hipMemcpyAsync(B.hv[B.cur] + o, hv, bytes, hipMemcpyHostToDevice, B.stream));
    HIP_TRY(hipMemcpyAsync(B.b + o, b, bytes, hipMemcpyHostToDevice, B.stream));
    HIP_TRY(hipStreamSynchronize(B.stream));
    B.uploaded[s] = 1;
    B.flags_valid = false;
    return SWE_OK;
}

int batch_download(Batch &B, int32_t s, double *eta, double *hu, double *hv, double *b) {
    if (s < 0 || s >= B.nscen) return fail(SWE_ERR_ARG, "scenario %d out of range", s);
    if (!B.uploaded[s]) return fail(SWE_ERR_STATE, "download: scenario %d has no state", s);
    const size_t bytes = sizeof(double) * (size_t)B.scen[s].n;
    const int64_t o = B.scen[s].cell_off;
    if (eta) HIP_TRY(hipMemcpyAsync(eta, B.eta[B.cur] + o, bytes, hipMemcpyDeviceToHost, B.stream));
    if (hu) HIP_TRY(hipMemcpyAsync(hu, B.hu[B.cur] + o, bytes, hipMemcpyDeviceToHost, B.stream));
    if (hv) HIP_TRY(hipMemcpyAsync(hv, B.hv[B.cur] + o, bytes, hipMemcpyDeviceToHost, B.stream));
    if (b) HIP_TRY(hipMemcpyAsync(b, B.b + o, bytes, hipMemcpyDeviceToHost, B.stream));
    HIP_TRY(hipStreamSynchronize(B.stream));
    return SWE_OK;
}

// One traversal with host-chosen dt per scenario; results per scenario.
int batch_step(Batch &B, const double *dt, int forward, double *dt_cfl, int64_t *n_refine,
               int8_t *refine_host) {
    if (!dt || !dt_cfl) return fail(SWE_ERR_ARG, "step: null argument");
    if (!all_uploaded(B)) return fail(SWE_ERR_STATE, "step: not every scenario has a state");
    HIP_TRY(hipMemcpyAsync(B.d_dt, dt, sizeof(double) * (size_t)B.nscen, hipMemcpyHostToDevice,
                           B.stream));
    swe::StepArgs a = make_args(B, B.cur, 1 - B.cur, forward);
    a.dt_host = B.d_dt;
    a.refine = (refine_host || B.keep_flags) ? B.refine : nullptr;
    HIP_TRY(swe::launch_step(a, B.solver, B.chunk, B.grid, B.stream));
    std::vector<swe::Slots> res((size_t)B.nscen);
    HIP_TRY(hipMemcpyAsync(res.data(), B.d_slots, sizeof(swe::Slots) * res.size(),
                           hipMemcpyDeviceToHost, B.stream));
    if (refine_host)
        HIP_TRY(hipMemcpyAsync(refine_host, B.refine, (size_t)B.ncells, hipMemcpyDeviceToHost,
                               B.stream));
    HIP_TRY(hipStreamSynchronize(B.stream));
    for (int32_t s = 0; s < B.nscen; ++s) {
        dt_cfl[s] = res[s].res_dtcfl;
        if (n_refine) n_refine[s] = (int64_t)res[s].res_nref;
    }
    B.cur = 1 - B.cur;
    B.flags_valid = a.refine != nullptr;
    return SWE_OK;
}

// n_steps traversals with the dt chain on the device, per scenario.
int batch_run(Batch &B, int64_t n_steps, const double *dt0, int forward, double dt_clamp,
              double *time, double *dt_next) {
    // dt0 == NULL: continue the device dt chain where the last run left it;
    // time == dt_next == NULL: do not wait for the steps
    if (n_steps < 0 || (dt0 && !time)) return fail(SWE_ERR_ARG, "run: invalid argument");
    if (!all_uploaded(B)) return fail(SWE_ERR_STATE, "run: not every scenario has a state");
    bool table_changed = false;
    for (auto &s : B.scen)
        if (s.dt_clamp != dt_clamp) {
            s.dt_clamp = dt_clamp;
            table_changed = true;
        }
    if (table_changed)
        HIP_TRY(hipMemcpyAsync(B.d_scen, B.scen.data(), sizeof(swe::Scen) * B.scen.size(),
                               hipMemcpyHostToDevice, B.stream));
    // the chain starts from dt0 (clamped as pre_traversal_grid_op does) and the
    // caller's time
    if (dt0) {
        std::vector<swe::Slots> init((size_t)B.nscen);
        for (int32_t s = 0; s < B.nscen; ++s) {
            init[s].dt = (dt_clamp > 0.0 && dt_clamp < dt0[s]) ? dt_clamp : dt0[s];
            init[s].time = time[s];
        }
        HIP_TRY(hipMemcpyAsync(B.d_slots, init.data(), sizeof(swe::Slots) * init.size(),
                               hipMemcpyHostToDevice, B.stream));
    }

    auto one_step = [&](hipStream_t st, int from, int fwd) -> hipError_t {
        swe::StepArgs a = make_args(B, from, 1 - from, fwd);
        a.device_dt = 1;
        a.refine = B.keep_flags ? B.refine : nullptr;
        return swe::launch_step(a, B.solver, B.chunk, B.grid, st);
    };
    int fwd = forward ? 1 : 0;
    const int64_t pairs = n_steps / 2;
    if (pairs > 0) {
        GraphCache &gc = B.graph[B.cur][fwd];
        if (!gc.exec || gc.cur != B.cur || gc.forward != fwd || gc.chunk != B.chunk ||
            gc.grid != B.grid || gc.keep != B.keep_flags || gc.jumps != B.jump_flags ||
            gc.tol != B.tol ||
            gc.jump_refine != B.jump_refine || gc.jump_coarsen != B.jump_coarsen) {
            if (gc.exec) {
                (void)hipGraphExecDestroy(gc.exec);
                gc.exec = nullptr;
            }
            // capture on the library's own stream, replay on B.stream
            HIP_TRY(hipStreamSynchronize(B.stream));
            hipGraph_t graph;
            HIP_TRY(hipStreamBeginCapture(B.own_stream, hipStreamCaptureModeThreadLocal));
            hipError_t e1 = one_step(B.own_stream, B.cur, fwd);
            hipError_t e2 = e1 == hipSuccess ? one_step(B.own_stream, 1 - B.cur, 1 - fwd) : e1;
            hipError_t e3 = hipStreamEndCapture(B.own_stream, &graph);
            if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess)
                return fail(SWE_ERR_HIP, "run: graph capture failed: %s",
                            hipGetErrorString(e1 != hipSuccess ? e1 : (e2 != hipSuccess ? e2 : e3)));
            hipError_t e4 = hipGraphInstantiate(&gc.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e4 != hipSuccess) return fail(SWE_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e4));
            gc.cur = B.cur;
            gc.forward = fwd;
            gc.chunk = B.chunk;
            gc.grid = B.grid;
            gc.keep = B.keep_flags;
            gc.jumps = B.jump_flags;
            gc.tol = B.tol;
            gc.jump_refine = B.jump_refine;
            gc.jump_coarsen = B.jump_coarsen;
        }
        for (int64_t i = 0; i < pairs; ++i) HIP_TRY(hipGraphLaunch(gc.exec, B.stream));
    }
    if (n_steps % 2) {
        HIP_TRY(one_step(B.stream, B.cur, fwd));
        B.cur = 1 - B.cur;
    }
    if (n_steps > 0) B.flags_valid = B.keep_flags != 0;
    if (!time && !dt_next) return SWE_OK;
    std::vector<swe::Slots> res((size_t)B.nscen);
    HIP_TRY(hipMemcpyAsync(res.data(), B.d_slots, sizeof(swe::Slots) * res.size(),
                           hipMemcpyDeviceToHost, B.stream));
    HIP_TRY(hipStreamSynchronize(B.stream));
    for (int32_t s = 0; s < B.nscen; ++s) {
        if (time) time[s] = res[s].time;
        if (dt_next) dt_next[s] = res[s].dt;
    }
    return SWE_OK;
}

template <typename T>
int dev_alloc(T **p, int64_t count) {
    *p = nullptr;
    if (hipMalloc(p, sizeof(T) * (size_t)(count > 0 ? count : 1)) != hipSuccess) {
        *p = nullptr;
        return fail(SWE_ERR_NOMEM, "device allocation of %lld elements failed", (long long)count);
    }
    return SWE_OK;
}

// Scratch of the remesher for n cells, with room to grow.
int ensure_adapt_scratch(Batch &B, int64_t n) {
    if (n <= B.ad_cap) return SWE_OK;
    const int64_t cap = n + n / 4;
    HIP_TRY(hipStreamSynchronize(B.stream));  // queued work may still use the old scratch
    (void)hipFree(B.ad_slots);
    (void)hipFree(B.ad_status);
    (void)hipFree(B.ad_off);
    (void)hipFree(B.ad_off2);
    (void)hipFree(B.ad_base);
    (void)hipFree(B.ad_bsum);
    (void)hipFree(B.ad_code);
    (void)hipFree(B.ad_kidx);
    (void)hipFree(B.ad_list);
    (void)hipFree(B.ad_eslot);
    B.ad_kidx = nullptr;
    B.ad_list = nullptr;
    B.ad_slots = B.ad_status = nullptr;
    B.ad_off = B.ad_off2 = B.ad_base = B.ad_bsum = nullptr;
    B.ad_off_cap = B.ad_off2_cap = 0;
    B.ad_code = nullptr;
    B.ad_eslot = nullptr;
    B.ad_cap = 0;
    B.ad_off_valid = false;
    SWE_TRY(dev_alloc(&B.ad_slots, cap));
    SWE_TRY(dev_alloc(&B.ad_status, 4 + cap / 256 + 1));  // status words, then sweep stamps
    HIP_TRY(hipMemsetAsync(B.ad_status, 0, sizeof(uint32_t) * (size_t)(4 + cap / 256 + 1),
                           B.stream));
    SWE_TRY(dev_alloc(&B.ad_off, cap));
    SWE_TRY(dev_alloc(&B.ad_off2, cap));
    B.ad_off_cap = B.ad_off2_cap = cap;
    SWE_TRY(dev_alloc(&B.ad_base, cap));
    SWE_TRY(dev_alloc(&B.ad_bsum, swe::adapt_scan_tiles(cap) + 1 + 8));
    SWE_TRY(dev_alloc(&B.ad_code, cap));
    SWE_TRY(dev_alloc(&B.ad_kidx, cap));
    SWE_TRY(dev_alloc(&B.ad_list, cap));
    SWE_TRY(dev_alloc(&B.ad_eslot, 6 * cap));
    B.ad_cap = cap;
    return SWE_OK;
}

// One remeshing step of a single-mesh handle: conformity of the refinement
// requests (the last step's flags on the device, or the caller's), split and
// merge in curve order, state transfer; the handle then steps the new mesh.
int batch_adapt(Batch &B, const int8_t *refine_host, int64_t *n_out, int *sweeps_out) {
    if (B.nscen != 1) return fail(SWE_ERR_ARG, "adapt: only a single-mesh handle remeshes");
    if (!all_uploaded(B)) return fail(SWE_ERR_STATE, "adapt: no state uploaded");
    const int64_t n = B.ncells;
    if (refine_host) {
        HIP_TRY(hipMemcpyAsync(B.refine, refine_host, (size_t)n, hipMemcpyHostToDevice, B.stream));
    } else if (!B.flags_valid) {
        return fail(SWE_ERR_STATE,
                    "adapt: no refinement flags on the device (pass them, or keep the steps' "
                    "flags with swe_keep_refinement_flags)");
    }
    SWE_TRY(ensure_adapt_scratch(B, n));
    if (!B.ad_host) {
        // coherent: the device stores the results there while the host polls
        if (hipHostMalloc(&B.ad_host, 8 * sizeof(int64_t), hipHostMallocCoherent) != hipSuccess) {
            B.ad_host = nullptr;
            return fail(SWE_ERR_NOMEM, "adapt: pinned host allocation failed");
        }
        std::memset(B.ad_host, 0, 8 * sizeof(int64_t));
        if (hipHostGetDevicePointer((void **)&B.ad_host_dev, B.ad_host, 0) != hipSuccess) {
            (void)hipHostFree(B.ad_host);
            B.ad_host = nullptr;
            return fail(SWE_ERR_HIP, "adapt: no device address for the pinned results");
        }
    }
    const swe::Scen &sc = B.scen[0];
    swe::AdaptArgs a{};
    a.n = n;
    a.depth = B.depth;
    a.plotter = B.plotter;
    a.nbr = B.nbr;
    a.flags = B.refine;
    a.eta = B.eta[B.cur];
    a.hu = B.hu[B.cur];
    a.hv = B.hv[B.cur];
    a.b = B.b;
    a.dmin = sc.dmin;
    a.dmax = sc.dmax;
    a.slots = B.ad_slots;
    a.off = B.ad_off;
    a.base = B.ad_base;
    a.bsum = B.ad_bsum;
    a.code = B.ad_code;
    a.kidx = B.ad_kidx;
    a.list = B.ad_list;
    a.eslot = B.ad_eslot;
    a.status = B.ad_status;
    a.stamp = B.ad_status + 4;
    a.raster = B.raster;
    a.rnx = B.rnx;
    a.rny = B.rny;
    a.rx0 = B.rx0;
    a.ry0 = B.ry0;
    a.rdx = B.rdx;
    a.rdy = B.rdy;
    a.scaling = sc.scaling;
    a.offset_x = B.roffset_x;
    a.offset_y = B.roffset_y;
    if (B.raster) {
        // SWE_initialize.f90:127-130: the depth at which a cell is as large as
        // a grid cell, nint(log(scaling^2 / (dx dy)) / log(2))
        const long data_depth = std::lround(std::log(sc.scaling * sc.scaling / (B.rdx * B.rdy)) /
                                            std::log(2.0));
        a.raster_cap = (int)std::min<long>(sc.dmax + 3, data_depth + 3);
    }
    const uint32_t epoch = ++B.ad_epoch;
    if (swe::stamp_tag_wraps(epoch)) {
        // the sweeps' 16-bit call tag starts again at 0: the stamps and the
        // last-change word of the calls 65536 back would read as this call's
        HIP_TRY(hipMemsetAsync(B.ad_status + 1, 0, sizeof(uint32_t), B.stream));
        HIP_TRY(hipMemsetAsync(a.stamp, 0, sizeof(uint32_t) * (size_t)(B.ad_cap / 256 + 1),
                               B.stream));
    }
    swe::AdaptPlan plan{};
    HIP_TRY(swe::adapt_plan(a, B.stream, B.ad_last_sweeps, B.ad_off_valid, B.ad_last_changed,
                            1 << 16, epoch, B.ad_host, B.ad_host_dev, &B.ad_seq, &plan));
    B.ad_last_changed = (plan.status & 2u) != 0;
    const int sweeps = plan.sweeps;
    if (sweeps < 0) return fail(SWE_ERR_STATE, "adapt: the conformity sweeps did not converge");
    // the next call starts with as many sweeps as changed a word this time
    B.ad_last_sweeps = plan.last_change > 1 ? plan.last_change : 1;
    const int64_t n_new = plan.n_new, tree = plan.tree_total;
    const uint32_t status = plan.status & 1u;
    if (tree != (int64_t(1) << 61))
        return fail(SWE_ERR_ARG,
                    "adapt: the mesh is not a leaf sequence of the two root triangles (its depths "
                    "cover %.17g of the unit square)", std::ldexp((double)tree, -61));
    if (status) return fail(SWE_ERR_STATE, "adapt: inconsistent coarsening (status %u)", status);
    B.ad_off_valid = true;
    if (n_new <= 0 || n_new > (int64_t(1) << 29))
        return fail(SWE_ERR_ARG, "adapt: the new mesh would have %lld cells (1..2^29)",
                    (long long)n_new);
    B.ad_record_n = n;
    B.ad_record_sampled = B.raster != nullptr;
    if (!(plan.status & 2u)) {  // every cell keeps its place: the mesh and state stand
        B.ad_record_n = -n;     // an identity record
        B.flags_valid = false;
        if (n_out) *n_out = n;
        if (sweeps_out) *sweeps_out = sweeps;
        return SWE_OK;
    }
    B.ad_off_valid = false;

    // outputs: the spare set when the new mesh fits, else fresh buffers with
    // room to grow (everything is allocated before anything is freed)
    const bool fresh = n_new > B.cap;
    const int64_t cap = fresh ? std::min<int64_t>(n_new + n_new / 4, int64_t(1) << 29) : B.cap;
    double *eo = nullptr, *ho = nullptr, *vo = nullptr, *bo = nullptr;
    double *e1 = nullptr, *h1 = nullptr, *v1 = nullptr;
    int32_t *no = nullptr;
    int8_t *po = nullptr, *dpo = nullptr, *ro = nullptr;
    auto drop = [&]() {
        for (void *p : {(void *)eo, (void *)ho, (void *)vo, (void *)bo, (void *)e1, (void *)h1,
                        (void *)v1, (void *)no, (void *)po, (void *)dpo, (void *)ro})
            (void)hipFree(p);
    };
    if (fresh) {
        int rc = SWE_OK;
        if (rc == SWE_OK) rc = dev_alloc(&eo, cap);
        if (rc == SWE_OK) rc = dev_alloc(&ho, cap);
        if (rc == SWE_OK) rc = dev_alloc(&vo, cap);
        if (rc == SWE_OK) rc = dev_alloc(&bo, cap);
        if (rc == SWE_OK) rc = dev_alloc(&e1, cap);
        if (rc == SWE_OK) rc = dev_alloc(&h1, cap);
        if (rc == SWE_OK) rc = dev_alloc(&v1, cap);
        if (rc == SWE_OK) rc = dev_alloc(&no, 3 * cap);
        if (rc == SWE_OK) rc = dev_alloc(&po, cap);
        if (rc == SWE_OK) rc = dev_alloc(&dpo, cap);
        if (rc == SWE_OK) rc = dev_alloc(&ro, cap);
        if (rc != SWE_OK) {
            drop();
            return rc;
        }
    } else {
        if (B.cap2 < B.cap) {
            (void)hipFree(B.b2);
            (void)hipFree(B.nbr2);
            (void)hipFree(B.plotter2);
            (void)hipFree(B.depth2);
            B.b2 = nullptr;
            B.nbr2 = nullptr;
            B.plotter2 = B.depth2 = nullptr;
            B.cap2 = 0;
            SWE_TRY(dev_alloc(&B.b2, B.cap));
            SWE_TRY(dev_alloc(&B.nbr2, 3 * B.cap));
            SWE_TRY(dev_alloc(&B.plotter2, B.cap));
            SWE_TRY(dev_alloc(&B.depth2, B.cap));
            B.cap2 = B.cap;
        }
    }
    if (n_new > B.ad_off2_cap) {  // room for the new mesh's tree offsets
        HIP_TRY(hipStreamSynchronize(B.stream));
        (void)hipFree(B.ad_off2);
        B.ad_off2 = nullptr;
        B.ad_off2_cap = 0;
        SWE_TRY(dev_alloc(&B.ad_off2, n_new + n_new / 4));
        B.ad_off2_cap = n_new + n_new / 4;
    }
    a.off_out = B.ad_off2;
    a.n_changing = plan.n_changing;
    a.eta_out = fresh ? eo : B.eta[1 - B.cur];
    a.hu_out = fresh ? ho : B.hu[1 - B.cur];
    a.hv_out = fresh ? vo : B.hv[1 - B.cur];
    a.b_out = fresh ? bo : B.b2;
    a.nbr_out = fresh ? no : B.nbr2;
    a.plotter_out = fresh ? po : B.plotter2;
    a.depth_out = fresh ? dpo : B.depth2;
    // the new mesh's table entry, as rebuild_chunks lays it out for one mesh
    swe::Scen sn = B.scen[0];
    sn.n = n_new;
    sn.depth = -1;
    sn.chunk_off = 0;
    sn.nchunks = (n_new + B.chunk - 1) / B.chunk;
    a.scen_dev = B.d_scen;
    a.scen_new = sn;
    hipError_t e = swe::adapt_emit(a, B.stream, plan.prefix_done);
    if (e != hipSuccess) {
        if (fresh) drop();
        return fail(SWE_ERR_HIP, "adapt: %s", hipGetErrorString(e));
    }
    if (fresh) {
        // the emit kernels read the old buffers
        HIP_TRY(hipStreamSynchronize(B.stream));
        for (int i = 0; i < 2; ++i) {
            (void)hipFree(B.eta[i]);
            (void)hipFree(B.hu[i]);
            (void)hipFree(B.hv[i]);
        }
        for (void *p : {(void *)B.b, (void *)B.nbr, (void *)B.plotter, (void *)B.depth,
                        (void *)B.refine, (void *)B.b2, (void *)B.nbr2, (void *)B.plotter2,
                        (void *)B.depth2})
            (void)hipFree(p);
        B.b2 = nullptr;
        B.nbr2 = nullptr;
        B.plotter2 = B.depth2 = nullptr;
        B.cap2 = 0;
        B.eta[0] = eo;
        B.hu[0] = ho;
        B.hv[0] = vo;
        B.eta[1] = e1;
        B.hu[1] = h1;
        B.hv[1] = v1;
        B.b = bo;
        B.nbr = no;
        B.plotter = po;
        B.depth = dpo;
        B.refine = ro;
        B.cur = 0;
        B.cap = cap;
    } else {
        std::swap(B.b, B.b2);
        std::swap(B.nbr, B.nbr2);
        std::swap(B.plotter, B.plotter2);
        std::swap(B.depth, B.depth2);
        B.cur = 1 - B.cur;
    }
    std::swap(B.ad_off, B.ad_off2);  // the new mesh's offsets stay valid
    std::swap(B.ad_off_cap, B.ad_off2_cap);
    B.ad_off_valid = true;
    B.ncells = B.ntopo = n_new;
    B.scen[0] = sn;
    B.uniform_depth[0] = -1;
    B.flags_valid = false;
    SWE_TRY(rebuild_chunks(B, false));
    if (n_out) *n_out = n_new;
    if (sweeps_out) *sweeps_out = sweeps;
    return SWE_OK;
}

// Forward traversal on caller-owned device buffers (all scenarios), dt read
// from device memory per scenario, limiter masks recorded (nullable).
int batch_forward_device(Batch &B, const double *eta, const double *hu, const double *hv,
                         const double *b, double *eta_out, double *hu_out, double *hv_out,
                         const double *dt_dev, int forward, double *dt_cfl_dev, int8_t *masks,
                         void *stream) {
    if (!eta || !hu || !hv || !b || !eta_out || !hu_out || !hv_out || !dt_dev)
        return fail(SWE_ERR_ARG, "forward_device: null argument");
    if (eta_out == eta || hu_out == hu || hv_out == hv)
        return fail(SWE_ERR_ARG, "forward_device: outputs must not alias inputs");
    hipStream_t st = stream ? (hipStream_t)stream : B.stream;
    swe::StepArgs a = make_args(B, 0, 1, forward);
    a.eta = eta;
    a.hu = hu;
    a.hv = hv;
    a.b = b;
    a.eta_out = eta_out;
    a.hu_out = hu_out;
    a.hv_out = hv_out;
    a.dt_host = dt_dev;
    a.masks = masks;
    a.out_dtcfl = dt_cfl_dev;
    a.skip_finalize = dt_cfl_dev ? 0 : 1;  // dt comes from the caller: no result is read
    a.tol = 0;  // the adjoint's sweeps: the rounding its kernels differentiate
    HIP_TRY(swe::launch_step(a, B.solver, B.chunk, B.grid, st));
    return SWE_OK;
}

int batch_adjoint_device(Batch &B, const double *eta, const double *hu, const double *hv,
                         const double *b, const int8_t *masks, const double *leta,
                         const double *lhu, const double *lhv, double *leta_out, double *lhu_out,
                         double *lhv_out, double *gb, const double *dt_dev, int forward,
                         void *stream) {
    if (!eta || !hu || !hv || !b || !masks || !leta || !lhu || !lhv || !leta_out || !lhu_out ||
        !lhv_out || !gb || !dt_dev)
        return fail(SWE_ERR_ARG, "adjoint_device: null argument");
    if (leta_out == leta || lhu_out == lhu || lhv_out == lhv)
        return fail(SWE_ERR_ARG, "adjoint_device: outputs must not alias inputs");
    hipStream_t st = stream ? (hipStream_t)stream : B.stream;
    swe::AdjArgs a{};
    a.eta = eta;
    a.hu = hu;
    a.hv = hv;
    a.b = b;
    a.masks = masks;
    a.leta = leta;
    a.lhu = lhu;
    a.lhv = lhv;
    a.leta_out = leta_out;
    a.lhu_out = lhu_out;
    a.lhv_out = lhv_out;
    a.gb = gb;
    a.nbr = B.nbr;
    a.plotter = B.plotter;
    a.depth = B.depth;
    a.scen = B.d_scen;
    a.chunk_scen = B.d_chunk_scen;
    a.nchunks = B.nchunks;
    a.dt_dev = dt_dev;
    a.g = B.g;
    a.inv16g = 1.0 / (16.0 * B.g);
    a.bc = B.bc;
    a.forward = forward ? 1 : 0;
    if (B.solver != SWE_SOLVER_AUG_RIEMANN && B.solver != SWE_SOLVER_FWAVE)
        return fail(SWE_ERR_ARG, "adjoint: the augmented and f-wave solvers only (solver %d)",
                    B.solver);
    HIP_TRY(swe::launch_adjoint(a, B.solver, B.chunk, B.grid, st));
    return SWE_OK;
}

int batch_set_stream(Batch &B, void *stream) {
    HIP_TRY(hipStreamSynchronize(B.stream));
    B.stream = stream ? (hipStream_t)stream : B.own_stream;
    return SWE_OK;
}

int batch_set_arithmetic(Batch &B, int mode) {
    if (mode != SWE_ARITH_EXACT && mode != SWE_ARITH_TOLERANCED)
        return fail(SWE_ERR_ARG, "arithmetic mode %d (SWE_ARITH_EXACT or SWE_ARITH_TOLERANCED)",
                    mode);
    HIP_TRY(hipStreamSynchronize(B.stream));
    B.tol = mode == SWE_ARITH_TOLERANCED ? 1 : 0;
    return SWE_OK;
}

int batch_set_tuning(Batch &B, int chunk, int64_t grid) {
    bool known = false;
    for (int c : swe::CHUNK_SIZES) known = known || c == chunk;
    if (!known) return fail(SWE_ERR_ARG, "chunk size %d not compiled (128, 256 or 512)", chunk);
    if (grid < 0 || grid % 8) return fail(SWE_ERR_ARG, "grid %lld must be 0 or a multiple of 8", (long long)grid);
    HIP_TRY(hipStreamSynchronize(B.stream));
    B.grid = grid;
    if (chunk != B.chunk) {
        B.chunk = chunk;
        SWE_TRY(rebuild_chunks(B));
    }
    return SWE_OK;
}

}  // namespace

extern "C" {

size_t swe_last_error(char *buf, size_t size) {
    if (buf && size) {
        std::strncpy(buf, g_last_error.c_str(), size - 1);
        buf[size - 1] = '\0';
    }
    return g_last_error.size();
}

int swe_chunk_cells(void) { return swe::DEFAULT_CHUNK; }

// ---- single mesh -----------------------------------------------------------

int swe_create(const swe_mesh_desc *desc, swe_handle **out) {
    if (!desc || !out) return fail(SWE_ERR_ARG, "swe_create: null argument");
    *out = nullptr;
    swe_handle *h = new swe_handle();
    int rc = batch_init(h->B, 1, desc);
    if (rc != SWE_OK) {
        std::string msg = "swe_create: " + g_last_error;
        release(h->B);
        delete h;
        return fail(rc, "%s", msg.c_str());
    }
    *out = h;
    return SWE_OK;
}

int swe_destroy(swe_handle *h) {
    if (!h) return SWE_OK;
    (void)hipStreamSynchronize(h->B.stream);
    release(h->B);
    delete h;
    return SWE_OK;
}

int swe_upload(swe_handle *h, const double *eta, const double *hu, const double *hv,
               const double *b) {
    if (!h) return fail(SWE_ERR_ARG, "swe_upload: null handle");
    return batch_upload(h->B, 0, eta, hu, hv, b);
}

int swe_download(swe_handle *h, double *eta, double *hu, double *hv, double *b) {
    if (!h) return fail(SWE_ERR_ARG, "swe_download: null handle");
    return batch_download(h->B, 0, eta, hu, hv, b);
}

int swe_step(swe_handle *h, double dt, int forward, double *dt_cfl, int64_t *n_refine,
             int8_t *refine) {
    if (!h) return fail(SWE_ERR_ARG, "swe_step: null handle");
    return batch_step(h->B, &dt, forward, dt_cfl, n_refine, refine);
}

int swe_run(swe_handle *h, int64_t n_steps, double dt0, int forward, double dt_clamp,
            double *time, double *dt_next) {
    if (!h) return fail(SWE_ERR_ARG, "swe_run: null handle");
    if (dt0 > 0.0 && !time) return fail(SWE_ERR_ARG, "swe_run: time is required with dt0 > 0");
    return batch_run(h->B, n_steps, dt0 > 0.0 ? &dt0 : nullptr, forward, dt_clamp, time, dt_next);
}

int64_t swe_cells(swe_handle *h) { return h ? h->B.ncells : -1; }

int swe_keep_refinement_flags(swe_handle *h, int on) {
    if (!h) return fail(SWE_ERR_ARG, "swe_keep_refinement_flags: null handle");
    h->B.keep_flags = on ? 1 : 0;
    h->B.jump_flags = 0;
    return SWE_OK;
}

int swe_keep_eta_jump_flags(swe_handle *h, double refine_tol, double coarsen_tol) {
    if (!h) return fail(SWE_ERR_ARG, "swe_keep_eta_jump_flags: null handle");
    if (h->B.nscen != 1) return fail(SWE_ERR_ARG, "swe_keep_eta_jump_flags: single-mesh handles only");
    if (!(refine_tol >= 0.0) || !(coarsen_tol >= 0.0))
        return fail(SWE_ERR_ARG, "swe_keep_eta_jump_flags: tolerances must be >= 0");
    h->B.keep_flags = 1;
    h->B.jump_flags = 1;
    h->B.jump_refine = refine_tol;
    h->B.jump_coarsen = coarsen_tol;
    return SWE_OK;
}

int swe_adapt(swe_handle *h, const int8_t *refine, int64_t *n_cells, int *sweeps) {
    if (!h) return fail(SWE_ERR_ARG, "swe_adapt: null handle");
    return batch_adapt(h->B, refine, n_cells, sweeps);
}

int swe_set_bathymetry_raster(swe_handle *h, int64_t nx, int64_t ny, double x0, double y0,
                              double dx, double dy, double offset_x, double offset_y,
                              const double *values) {
    if (!h) return fail(SWE_ERR_ARG, "swe_set_bathymetry_raster: null handle");
    Batch &B = h->B;
    if (!values) {
        (void)hipFree(B.raster);
        B.raster = nullptr;
        return SWE_OK;
    }
    if (nx < 2 || ny < 2 || nx > (int64_t(1) << 20) || ny > (int64_t(1) << 20) ||
        !(dx > 0.0) || !(dy > 0.0) || !std::isfinite(x0) || !std::isfinite(y0) ||
        !std::isfinite(offset_x) || !std::isfinite(offset_y))
        return fail(SWE_ERR_ARG,
                    "swe_set_bathymetry_raster: need 2 <= nx, ny <= 2^20 and dx, dy > 0");
    double *d = nullptr;
    SWE_TRY(dev_alloc(&d, nx * ny));
    hipError_t e = hipMemcpy(d, values, sizeof(double) * (size_t)(nx * ny), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d);
        return fail(SWE_ERR_HIP, "swe_set_bathymetry_raster: %s", hipGetErrorString(e));
    }
    (void)hipFree(B.raster);
    B.raster = d;
    B.rnx = nx;
    B.rny = ny;
    B.rx0 = x0;
    B.ry0 = y0;
    B.rdx = dx;
    B.rdy = dy;
    B.roffset_x = offset_x;
    B.roffset_y = offset_y;
    return SWE_OK;
}

int swe_flag_eta_jumps(swe_handle *h, double refine_tol, double coarsen_tol) {
    if (!h) return fail(SWE_ERR_ARG, "swe_flag_eta_jumps: null handle");
    Batch &B = h->B;
    if (B.nscen != 1) return fail(SWE_ERR_ARG, "swe_flag_eta_jumps: single-mesh handles only");
    if (!all_uploaded(B)) return fail(SWE_ERR_STATE, "swe_flag_eta_jumps: no state uploaded");
    if (!(refine_tol >= 0.0) || !(coarsen_tol >= 0.0))
        return fail(SWE_ERR_ARG, "swe_flag_eta_jumps: tolerances must be >= 0");
    const swe::Scen &sc = B.scen[0];
    HIP_TRY(swe::flag_eta_jumps(B.ncells, B.eta[B.cur], B.b, B.nbr, B.depth, sc.drytol,
                                refine_tol, coarsen_tol, sc.dmin, sc.dmax, B.refine, B.stream));
    B.flags_valid = true;
    return SWE_OK;
}

int swe_adapt_record_device(swe_handle *h, int8_t *code, int64_t *base, int64_t *n_old,
                            int *b_sampled, void *stream) {
    if (!h || !n_old) return fail(SWE_ERR_ARG, "swe_adapt_record_device: null argument");
    Batch &B = h->B;
    if (B.ad_record_n == -1)
        return fail(SWE_ERR_STATE, "swe_adapt_record_device: no remeshing yet");
    const int64_t n = B.ad_record_n < 0 ? -B.ad_record_n : B.ad_record_n;
    *n_old = n;
    if (b_sampled) *b_sampled = B.ad_record_sampled ? 1 : 0;
    if (!code && !base) return SWE_OK;
    if (!code || !base) return fail(SWE_ERR_ARG, "swe_adapt_record_device: code and base go together");
    hipStream_t st = stream ? (hipStream_t)stream : B.stream;
    if (B.ad_record_n < 0) {
        HIP_TRY(swe::identity_record(n, code, base, st));
    } else {
        HIP_TRY(hipMemcpyAsync(code, B.ad_code, (size_t)n, hipMemcpyDeviceToDevice, st));
        HIP_TRY(swe::record_base(n, B.ad_kidx, B.ad_base, base, st));
    }
    return SWE_OK;
}

int swe_transfer_adjoint_device(int64_t n_old, const int8_t *code, const int64_t *base,
                                int b_sampled, const double *leta, const double *lhu,
                                const double *lhv, const double *gb, double *leta_out,
                                double *lhu_out, double *lhv_out, double *gb_out, void *stream) {
    if (n_old <= 0 || !code || !base || !leta || !lhu || !lhv || !gb || !leta_out || !lhu_out ||
        !lhv_out || !gb_out)
        return fail(SWE_ERR_ARG, "swe_transfer_adjoint_device: invalid argument");
    HIP_TRY(swe::transfer_adjoint(n_old, code, base, b_sampled, leta, lhu, lhv, gb, leta_out,
                                  lhu_out, lhv_out, gb_out, (hipStream_t)stream));
    return SWE_OK;
}

int swe_download_mesh(swe_handle *h, int8_t *depth, int8_t *plotter, int32_t *nbr) {
    if (!h) return fail(SWE_ERR_ARG, "swe_download_mesh: null handle");
    Batch &B = h->B;
    if (B.nscen != 1) return fail(SWE_ERR_ARG, "swe_download_mesh: a batch has no single mesh");
    const int64_t n = B.ncells;
    if (depth) HIP_TRY(hipMemcpyAsync(depth, B.depth, (size_t)n, hipMemcpyDeviceToHost, B.stream));
    if (plotter)
        HIP_TRY(hipMemcpyAsync(plotter, B.plotter, (size_t)n, hipMemcpyDeviceToHost, B.stream));
    if (nbr)
        HIP_TRY(hipMemcpyAsync(nbr, B.nbr, sizeof(int32_t) * 3 * (size_t)n, hipMemcpyDeviceToHost,
                               B.stream));
    HIP_TRY(hipStreamSynchronize(B.stream));
    if (nbr)  // (o << 2) | back edge -> o
        for (int64_t i = 0; i < 3 * n; ++i) nbr[i] = nbr[i] < 0 ? -1 : (nbr[i] >> 2);
    return SWE_OK;
}

int swe_set_stream(swe_handle *h, void *stream) {
    if (!h) return fail(SWE_ERR_ARG, "swe_set_stream: null handle");
    return batch_set_stream(h->B, stream);
}

int swe_set_tuning(swe_handle *h, int chunk_cells, int64_t grid_blocks) {
    if (!h) return fail(SWE_ERR_ARG, "swe_set_tuning: null handle");
    return batch_set_tuning(h->B, chunk_cells, grid_blocks);
}

int swe_set_arithmetic(swe_handle *h, int mode) {
    if (!h) return fail(SWE_ERR_ARG, "swe_set_arithmetic: null handle");
    return batch_set_arithmetic(h->B, mode);
}

int swe_get_tuning(swe_handle *h, int *chunk_cells, int64_t *grid_blocks) {
    if (!h || !chunk_cells || !grid_blocks)
        return fail(SWE_ERR_ARG, "swe_get_tuning: null argument");
    *chunk_cells = h->B.chunk;
    *grid_blocks = h->B.grid;
    return SWE_OK;
}

int swe_step_device(swe_handle *h, const double *eta, const double *hu, const double *hv,
                    const double *b, double *eta_out, double *hu_out, double *hv_out, double dt,
                    int forward, double *dt_cfl_dev, int64_t *n_refine_dev, int8_t *refine_dev,
                    void *stream) {
    if (!h || !eta || !hu || !hv || !b || !eta_out || !hu_out || !hv_out || !dt_cfl_dev)
        return fail(SWE_ERR_ARG, "swe_step_device: null argument");
    if (eta_out == eta || hu_out == hu || hv_out == hv)
        return fail(SWE_ERR_ARG, "swe_step_device: outputs must not alias inputs");
    Batch &B = h->B;
    hipStream_t s = stream ? (hipStream_t)stream : B.stream;
    swe::StepArgs a = make_args(B, 0, 1, forward);
    a.eta = eta;
    a.hu = hu;
    a.hv = hv;
    a.b = b;
    a.eta_out = eta_out;
    a.hu_out = hu_out;
    a.hv_out = hv_out;
    a.dt = dt;
    a.refine = refine_dev;
    a.out_dtcfl = dt_cfl_dev;
    a.out_nref = n_refine_dev;
    HIP_TRY(swe::launch_step(a, B.solver, B.chunk, B.grid, s));
    return SWE_OK;
}

int swe_forward_device(swe_handle *h, const double *eta, const double *hu, const double *hv,
                       const double *b, double *eta_out, double *hu_out, double *hv_out,
                       const double *dt_dev, int forward, double *dt_cfl_dev, int8_t *masks_dev,
                       void *stream) {
    if (!h) return fail(SWE_ERR_ARG, "swe_forward_device: null handle");
    return batch_forward_device(h->B, eta, hu, hv, b, eta_out, hu_out, hv_out, dt_dev, forward,
                                dt_cfl_dev, masks_dev, stream);
}

int swe_adjoint_step_device(swe_handle *h, const double *eta, const double *hu, const double *hv,
                            const double *b, const int8_t *masks, const double *leta,
                            const double *lhu, const double *lhv, double *leta_out,
                            double *lhu_out, double *lhv_out, double *gb, const double *dt_dev,
                            int forward, void *stream) {
    if (!h) return fail(SWE_ERR_ARG, "swe_adjoint_step_device: null handle");
    return batch_adjoint_device(h->B, eta, hu, hv, b, masks, leta, lhu, lhv, leta_out, lhu_out,
                                lhv_out, gb, dt_dev, forward, stream);
}

}  // extern "C"

// ---- the step without device allocations -----------------------------------
//
// swe_step_device_noalloc keeps the step's scratch (the dt/result slots of the
// finalize kernel and the per-chunk CFL minima and refinement counts) in one
// file-scope device array, `scratch`, of a fixed size; every call on a device
// shares that device's copy.  Host code never passes the array's name to a
// kernel (on the host it names a stand-in, not device memory): the device
// address comes from hipGetSymbolAddress, once per device, and is kept in
// g_scratch.
namespace {
constexpr int64_t NOALLOC_MAX_CHUNKS = 65536;
constexpr int64_t NOALLOC_MAX_CELLS = NOALLOC_MAX_CHUNKS * swe::DEFAULT_CHUNK;  // 16,777,216
constexpr size_t NOALLOC_SLOT_BYTES = 256;  // one Slots record, padded
constexpr size_t NOALLOC_BYTES =
    NOALLOC_SLOT_BYTES + (size_t)NOALLOC_MAX_CHUNKS * (sizeof(double) + sizeof(int32_t));
static_assert(sizeof(swe::Slots) <= NOALLOC_SLOT_BYTES, "slots record");
}  // namespace

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace {
std::mutex g_scratch_mutex;
unsigned char *g_scratch[MAX_DEVICES] = {};  // device address of `scratch` per device

// The device address of `scratch` on the current device (looked up on the
// first call there).
int device_scratch(int device, unsigned char **out) {
    if (device < 0 || device >= MAX_DEVICES)
        return fail(SWE_ERR_ARG, "device %d beyond the %d the library serves", device, MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[device]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(SWE_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch");
        g_scratch[device] = (unsigned char *)p;
    }
    *out = g_scratch[device];
    return SWE_OK;
}
}  // namespace

extern "C" {

int swe_step_device_noalloc(const swe_device_mesh *m, const double *eta, const double *hu,
                            const double *hv, const double *b, double *eta_out, double *hu_out,
                            double *hv_out, double dt, int forward, double *dt_cfl_dev,
                            int64_t *n_refine_dev, int8_t *refine_dev, void *stream) {
    if (!m || !eta || !hu || !hv || !b || !eta_out || !hu_out || !hv_out || !dt_cfl_dev)
        return fail(SWE_ERR_ARG, "swe_step_device_noalloc: null argument");
    if (!m->depth || !m->plotter || !m->nbr)
        return fail(SWE_ERR_ARG, "swe_step_device_noalloc: null mesh array");
    if (eta_out == eta || hu_out == hu || hv_out == hv)
        return fail(SWE_ERR_ARG, "swe_step_device_noalloc: outputs must not alias inputs");
    const int64_t n = m->n_cells;
    if (n <= 0 || n > NOALLOC_MAX_CELLS)
        return fail(SWE_ERR_ARG,
                    "swe_step_device_noalloc: n_cells %lld out of range (1..%lld, the size of "
                    "its fixed scratch)", (long long)n, (long long)NOALLOC_MAX_CELLS);
    if (!(m->scaling > 0.0) || !(m->g >= swe::MIN_G && m->g <= swe::MAX_G) ||
        !(m->dry_tolerance >= swe::MIN_DRY_TOLERANCE))
        return fail(SWE_ERR_ARG,
                    "swe_step_device_noalloc: scaling must be > 0, g in [%g, %g], "
                    "dry_tolerance >= %g", swe::MIN_G, swe::MAX_G, swe::MIN_DRY_TOLERANCE);
    if (m->solver < SWE_SOLVER_AUG_RIEMANN || m->solver > SWE_SOLVER_LLF_BATH)
        return fail(SWE_ERR_ARG, "swe_step_device_noalloc: unknown solver %d", m->solver);
    if (m->boundary != SWE_BC_OUTFLOW && m->boundary != SWE_BC_WALL)
        return fail(SWE_ERR_ARG, "swe_step_device_noalloc: unknown boundary %d", m->boundary);
    if (m->max_depth < 0 || m->max_depth > 60 || m->min_depth > m->max_depth)
        return fail(SWE_ERR_ARG, "swe_step_device_noalloc: depths %d..%d out of range",
                    m->min_depth, m->max_depth);
    int device = 0;
    HIP_TRY(hipGetDevice(&device));
    SWE_TRY(ensure_tables(device));
    unsigned char *scr = nullptr;
    SWE_TRY(device_scratch(device, &scr));

    double leg[64];
    leg_table(leg);
    swe::StepArgs a{};
    a.eta = eta;
    a.hu = hu;
    a.hv = hv;
    a.b = b;
    a.eta_out = eta_out;
    a.hu_out = hu_out;
    a.hv_out = hv_out;
    a.nbr = m->nbr;
    a.plotter = m->plotter;
    a.depth = m->depth;
    a.refine = refine_dev;
    a.slots = (swe::Slots *)scr;
    a.part_min = (double *)(scr + NOALLOC_SLOT_BYTES);
    a.part_nref = (int32_t *)(scr + NOALLOC_SLOT_BYTES + sizeof(double) * NOALLOC_MAX_CHUNKS);
    a.out_dtcfl = dt_cfl_dev;
    a.out_nref = n_refine_dev;
    a.n_scen = 1;
    a.nchunks = (n + swe::DEFAULT_CHUNK - 1) / swe::DEFAULT_CHUNK;
    a.dt = dt;
    a.g = m->g;
    a.inv16g = 1.0 / (16.0 * m->g);
    a.bc = m->boundary;
    a.forward = forward ? 1 : 0;
    a.raw_nbr = 1;
    swe::Scen &sc = a.scen0;
    sc.n = n;
    sc.nchunks = a.nchunks;
    sc.scaling = m->scaling;
    sc.drytol = m->dry_tolerance;
    sc.refine_thr = (5.0 * m->scaling) * leg[m->max_depth + 1];  // volume_op :339
    sc.dt_clamp = -1.0;
    sc.depth = -1;
    sc.dmin = m->min_depth;
    sc.dmax = m->max_depth;
    HIP_TRY(swe::launch_step_raw(a, m->solver, (hipStream_t)stream));
    return SWE_OK;
}

// ---- batch of meshes -------------------------------------------------------

int swe_batch_create(int32_t n_scenarios, const swe_mesh_desc *descs, swe_batch **out) {
    if (!descs || !out || n_scenarios <= 0) return fail(SWE_ERR_ARG, "swe_batch_create: invalid argument");
    *out = nullptr;
    swe_batch *b = new swe_batch();
    int rc = batch_init(b->B, n_scenarios, descs);
    if (rc != SWE_OK) {
        std::string msg = "swe_batch_create: " + g_last_error;
        release(b->B);
        delete b;
        return fail(rc, "%s", msg.c_str());
    }
    *out = b;
    return SWE_OK;
}

int swe_batch_destroy(swe_batch *b) {
    if (!b) return SWE_OK;
    (void)hipStreamSynchronize(b->B.stream);
    release(b->B);
    delete b;
    return SWE_OK;
}

int swe_batch_upload(swe_batch *b, int32_t scenario, const double *eta, const double *hu,
                     const double *hv, const double *bathymetry) {
    if (!b) return fail(SWE_ERR_ARG, "swe_batch_upload: null batch");
    return batch_upload(b->B, scenario, eta, hu, hv, bathymetry);
}

int swe_batch_download(swe_batch *b, int32_t scenario, double *eta, double *hu, double *hv,
                       double *bathymetry) {
    if (!b) return fail(SWE_ERR_ARG, "swe_batch_download: null batch");
    return batch_download(b->B, scenario, eta, hu, hv, bathymetry);
}

int swe_batch_step(swe_batch *b, const double *dt, int forward, double *dt_cfl,
                   int64_t *n_refine) {
    if (!b) return fail(SWE_ERR_ARG, "swe_batch_step: null batch");
    return batch_step(b->B, dt, forward, dt_cfl, n_refine, nullptr);
}

int swe_batch_run(swe_batch *b, int64_t n_steps, const double *dt0, int forward,
                  double dt_clamp, double *time, double *dt_next) {
    if (!b) return fail(SWE_ERR_ARG, "swe_batch_run: null batch");
    return batch_run(b->B, n_steps, dt0, forward, dt_clamp, time, dt_next);
}

int swe_batch_set_tuning(swe_batch *b, int chunk_cells, int64_t grid_blocks) {
    if (!b) return fail(SWE_ERR_ARG, "swe_batch_set_tuning: null batch");
    return batch_set_tuning(b->B, chunk_cells, grid_blocks);
}

int swe_batch_set_arithmetic(swe_batch *b, int mode) {
    if (!b) return fail(SWE_ERR_ARG, "swe_batch_set_arithmetic: null batch");
    return batch_set_arithmetic(b->B, mode);
}

int swe_batch_get_tuning(swe_batch *b, int *chunk_cells, int64_t *grid_blocks) {
    if (!b || !chunk_cells || !grid_blocks) return fail(SWE_ERR_ARG, "swe_batch_get_tuning: null argument");
    *chunk_cells = b->B.chunk;
    *grid_blocks = b->B.grid;
    return SWE_OK;
}

int swe_batch_set_stream(swe_batch *b, void *stream) {
    if (!b) return fail(SWE_ERR_ARG, "swe_batch_set_stream: null batch");
    return batch_set_stream(b->B, stream);
}

int64_t swe_batch_cells(swe_batch *b) { return b ? b->B.ncells : -1; }

int swe_batch_forward_device(swe_batch *bt, const double *eta, const double *hu, const double *hv,
                             const double *b, double *eta_out, double *hu_out, double *hv_out,
                             const double *dt_dev, int forward, double *dt_cfl_dev,
                             int8_t *masks_dev, void *stream) {
    if (!bt) return fail(SWE_ERR_ARG, "swe_batch_forward_device: null batch");
    return batch_forward_device(bt->B, eta, hu, hv, b, eta_out, hu_out, hv_out, dt_dev, forward,
                                dt_cfl_dev, masks_dev, stream);
}

int swe_batch_adjoint_device(swe_batch *bt, const double *eta, const double *hu, const double *hv,
                             const double *b, const int8_t *masks, const double *leta,
                             const double *lhu, const double *lhv, double *leta_out,
                             double *lhu_out, double *lhv_out, double *gb, const double *dt_dev,
                             int forward, void *stream) {
    if (!bt) return fail(SWE_ERR_ARG, "swe_batch_adjoint_device: null batch");
    return batch_adjoint_device(bt->B, eta, hu, hv, b, masks, leta, lhu, lhv, leta_out, lhu_out,
                                lhv_out, gb, dt_dev, forward, stream);
}

int64_t swe_batch_offset(swe_batch *b, int32_t scenario) {
    if (!b || scenario < 0 || scenario >= b->B.nscen) return -1;
    return b->B.scen[scenario].cell_off;
}

}  // extern "C"
