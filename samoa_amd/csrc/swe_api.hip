// swe_api.hip -- host side of the C ABI declared in include/swe_mi355x.h.
//
// Owns the device buffers of one frozen mesh (SoA state double-buffered,
// packed topology), launches the fused step kernel, and keeps the reference's
// dt bookkeeping either on the host (swe_step) or on the device inside a
// captured HIP graph (swe_run).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "swe_internal.hpp"
#include "swe_mi355x.h"

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

std::mutex g_table_mutex;
std::vector<int> g_tables_on_device;

int ensure_tables(int device) {
    std::lock_guard<std::mutex> lock(g_table_mutex);
    for (int d : g_tables_on_device)
        if (d == device) return SWE_OK;
    double normal[17 * 3 * 2], leg[64];
    normal_table(normal);
    leg_table(leg);
    HIP_TRY(swe::upload_tables(normal, leg));
    g_tables_on_device.push_back(device);
    return SWE_OK;
}

// The captured pair of steps bakes in which buffer holds the state, the
// starting direction and the clamp; a pair returns the state to the buffer it
// started from, so the graph stays valid across swe_run calls.
struct GraphCache {
    hipGraphExec_t exec = nullptr;
    int cur = -1;
    int forward = -1;
    double dt_clamp = 0.0;
};

}  // namespace

struct swe_handle {
    int device = 0;
    int64_t n = 0;
    swe_mesh_desc desc{};
    double refine_thr = 0.0;
    hipStream_t stream = nullptr;         // where work is launched (caller may replace it)
    hipStream_t own_stream = nullptr;     // created by the library; used for graph capture
    double *eta[2] = {nullptr, nullptr}, *hu[2] = {nullptr, nullptr}, *hv[2] = {nullptr, nullptr};
    double *b = nullptr;
    int32_t *nbr = nullptr;
    int8_t *plotter = nullptr, *depth = nullptr, *refine = nullptr;
    swe::Slots *slots = nullptr;  // reduction slots and device dt chain
    int cur = 0;
    bool has_state = false;
    GraphCache graph;
};

namespace {

void release(swe_handle *h) {
    if (!h) return;
    // teardown: errors cannot be reported from here, the handle is gone either way
    if (h->graph.exec) (void)hipGraphExecDestroy(h->graph.exec);
    for (int i = 0; i < 2; ++i) {
        (void)hipFree(h->eta[i]);
        (void)hipFree(h->hu[i]);
        (void)hipFree(h->hv[i]);
    }
    (void)hipFree(h->b);
    (void)hipFree(h->nbr);
    (void)hipFree(h->plotter);
    (void)hipFree(h->depth);
    (void)hipFree(h->refine);
    (void)hipFree(h->slots);
    if (h->own_stream) (void)hipStreamDestroy(h->own_stream);
    delete h;
}

swe::StepArgs make_args(const swe_handle *h, const double *eta, const double *hu,
                        const double *hv, const double *b, double *eo, double *ho, double *vo,
                        double dt, int forward) {
    swe::StepArgs a{};
    a.eta = eta;
    a.hu = hu;
    a.hv = hv;
    a.b = b;
    a.eta_out = eo;
    a.hu_out = ho;
    a.hv_out = vo;
    a.nbr = h->nbr;
    a.plotter = h->plotter;
    a.depth = h->depth;
    a.n = h->n;
    a.dt = dt;
    a.scaling = h->desc.scaling;
    a.g = h->desc.g;
    a.drytol = h->desc.dry_tolerance;
    a.refine_thr = h->refine_thr;
    a.dmin = h->desc.min_depth;
    a.dmax = h->desc.max_depth;
    a.bc = h->desc.boundary;
    a.forward = forward ? 1 : 0;
    a.courant = h->desc.courant;
    a.slots = h->slots;
    return a;
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

int swe_chunk_cells(void) { return swe::CHUNK; }

int swe_create(const swe_mesh_desc *desc, swe_handle **out) {
    if (!desc || !out) return fail(SWE_ERR_ARG, "swe_create: null argument");
    *out = nullptr;
    const int64_t n = desc->n_cells;
    if (n <= 0 || n > (int64_t(1) << 29))
        return fail(SWE_ERR_ARG, "swe_create: n_cells %lld out of range (1..2^29)", (long long)n);
    if (!desc->depth || !desc->plotter || !desc->nbr)
        return fail(SWE_ERR_ARG, "swe_create: null mesh array");
    if (!(desc->scaling > 0.0) || !(desc->g > 0.0) || !(desc->dry_tolerance >= 0.0))
        return fail(SWE_ERR_ARG, "swe_create: scaling and g must be > 0, dry_tolerance >= 0");
    if (desc->solver != SWE_SOLVER_AUG_RIEMANN && desc->solver != SWE_SOLVER_FWAVE)
        return fail(SWE_ERR_ARG, "swe_create: unknown solver %d", desc->solver);
    if (desc->boundary != SWE_BC_OUTFLOW && desc->boundary != SWE_BC_WALL)
        return fail(SWE_ERR_ARG, "swe_create: unknown boundary %d", desc->boundary);
    if (desc->max_depth < 0 || desc->max_depth > 60)
        return fail(SWE_ERR_ARG, "swe_create: max_depth %d out of range", desc->max_depth);

    // validate the mesh and pack neighbours as (o << 2) | back edge
    double normal[17 * 3 * 2], leg[64];
    normal_table(normal);
    leg_table(leg);
    std::vector<int32_t> packed((size_t)(3 * n));
    for (int64_t c = 0; c < n; ++c) {
        const int d = desc->depth[c];
        const int p = desc->plotter[c];
        if (d < 0 || d > 60) return fail(SWE_ERR_ARG, "swe_create: depth[%lld] = %d", (long long)c, d);
        if (p < -8 || p > 8 || p == 0)
            return fail(SWE_ERR_ARG, "swe_create: plotter[%lld] = %d", (long long)c, p);
        for (int k = 0; k < 3; ++k) {
            const int32_t o = desc->nbr[3 * c + k];
            if (o < 0) {
                if (o != -1)
                    return fail(SWE_ERR_ARG, "swe_create: nbr[%lld] = %d", (long long)(3 * c + k), o);
                packed[3 * c + k] = -1;
                continue;
            }
            if (o >= n || o == c)
                return fail(SWE_ERR_ARG, "swe_create: nbr[%lld] = %d invalid", (long long)(3 * c + k), o);
            int ko = -1;
            for (int j = 0; j < 3; ++j)
                if (desc->nbr[3 * (int64_t)o + j] == (int32_t)c) ko = j;
            if (ko < 0)
                return fail(SWE_ERR_ARG, "swe_create: neighbour relation not symmetric at cell %lld edge %d",
                            (long long)c, k);
            const int po = desc->plotter[o];
            const double *nc = &normal[((p + 8) * 3 + k) * 2];
            const double *no = &normal[((po + 8) * 3 + ko) * 2];
            if (!(nc[0] == -no[0] && nc[1] == -no[1]))
                return fail(SWE_ERR_ARG, "swe_create: cells %lld and %d disagree on the normal of their shared edge",
                            (long long)c, o);
            packed[3 * c + k] = (int32_t)((o << 2) | ko);
        }
    }

    swe_handle *h = new swe_handle();
    h->n = n;
    h->desc = *desc;
    h->desc.depth = nullptr;
    h->desc.plotter = nullptr;
    h->desc.nbr = nullptr;
    h->refine_thr = (5.0 * desc->scaling) * leg[desc->max_depth + 1];
    hipError_t e = hipGetDevice(&h->device);
    int rc = SWE_OK;
    if (e == hipSuccess) rc = ensure_tables(h->device);
    if (e != hipSuccess || rc != SWE_OK) {
        release(h);
        return e != hipSuccess ? fail(SWE_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(e)) : rc;
    }
    const size_t bytes = sizeof(double) * (size_t)n;
    bool ok = hipStreamCreate(&h->own_stream) == hipSuccess;
    h->stream = h->own_stream;
    for (int i = 0; i < 2 && ok; ++i)
        ok = hipMalloc(&h->eta[i], bytes) == hipSuccess && hipMalloc(&h->hu[i], bytes) == hipSuccess &&
             hipMalloc(&h->hv[i], bytes) == hipSuccess;
    ok = ok && hipMalloc(&h->b, bytes) == hipSuccess &&
         hipMalloc(&h->nbr, sizeof(int32_t) * 3 * (size_t)n) == hipSuccess &&
         hipMalloc(&h->plotter, (size_t)n) == hipSuccess && hipMalloc(&h->depth, (size_t)n) == hipSuccess &&
         hipMalloc(&h->refine, (size_t)n) == hipSuccess &&
         hipMalloc(&h->slots, sizeof(swe::Slots)) == hipSuccess;
    if (!ok) {
        release(h);
        return fail(SWE_ERR_NOMEM, "swe_create: device allocation of %lld cells failed", (long long)n);
    }
    ok = hipMemcpy(h->nbr, packed.data(), sizeof(int32_t) * 3 * (size_t)n, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(h->plotter, desc->plotter, (size_t)n, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(h->depth, desc->depth, (size_t)n, hipMemcpyHostToDevice) == hipSuccess;
    swe::Slots init{};
    const double huge_sr = 1.7976931348623157e308;
    std::memcpy(&init.dtmin, &huge_sr, sizeof huge_sr);
    ok = ok && hipMemcpy(h->slots, &init, sizeof init, hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        release(h);
        return fail(SWE_ERR_HIP, "swe_create: topology upload failed");
    }
    *out = h;
    return SWE_OK;
}

int swe_destroy(swe_handle *h) {
    if (!h) return SWE_OK;
    (void)hipStreamSynchronize(h->stream);
    release(h);
    return SWE_OK;
}

int swe_upload(swe_handle *h, const double *eta, const double *hu, const double *hv,
               const double *b) {
    if (!h || !eta || !hu || !hv || !b) return fail(SWE_ERR_ARG, "swe_upload: null argument");
    const size_t bytes = sizeof(double) * (size_t)h->n;
    HIP_TRY(hipMemcpyAsync(h->eta[h->cur], eta, bytes, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->hu[h->cur], hu, bytes, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->hv[h->cur], hv, bytes, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->b, b, bytes, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->has_state = true;
    return SWE_OK;
}

int swe_download(swe_handle *h, double *eta, double *hu, double *hv, double *b) {
    if (!h) return fail(SWE_ERR_ARG, "swe_download: null handle");
    if (!h->has_state) return fail(SWE_ERR_STATE, "swe_download: no state uploaded");
    const size_t bytes = sizeof(double) * (size_t)h->n;
    if (eta) HIP_TRY(hipMemcpyAsync(eta, h->eta[h->cur], bytes, hipMemcpyDeviceToHost, h->stream));
    if (hu) HIP_TRY(hipMemcpyAsync(hu, h->hu[h->cur], bytes, hipMemcpyDeviceToHost, h->stream));
    if (hv) HIP_TRY(hipMemcpyAsync(hv, h->hv[h->cur], bytes, hipMemcpyDeviceToHost, h->stream));
    if (b) HIP_TRY(hipMemcpyAsync(b, h->b, bytes, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return SWE_OK;
}

int swe_step(swe_handle *h, double dt, int forward, double *dt_cfl, int64_t *n_refine,
             int8_t *refine) {
    if (!h || !dt_cfl) return fail(SWE_ERR_ARG, "swe_step: null argument");
    if (!h->has_state) return fail(SWE_ERR_STATE, "swe_step: no state uploaded");
    const int c = h->cur, o = 1 - h->cur;
    swe::StepArgs a = make_args(h, h->eta[c], h->hu[c], h->hv[c], h->b, h->eta[o], h->hu[o],
                                h->hv[o], dt, forward);
    a.refine = refine ? h->refine : nullptr;
    HIP_TRY(swe::launch_step(a, h->desc.solver, h->stream));
    swe::Slots res;
    HIP_TRY(hipMemcpyAsync(&res, h->slots, sizeof res, hipMemcpyDeviceToHost, h->stream));
    if (refine)
        HIP_TRY(hipMemcpyAsync(refine, h->refine, (size_t)h->n, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    *dt_cfl = res.res_dtcfl;
    if (n_refine) *n_refine = (int64_t)res.res_nref;
    h->cur = o;
    return SWE_OK;
}

int swe_run(swe_handle *h, int64_t n_steps, double dt0, int forward, double dt_clamp,
            double *time_added, double *dt_next) {
    if (!h || n_steps < 0) return fail(SWE_ERR_ARG, "swe_run: invalid argument");
    if (!h->has_state) return fail(SWE_ERR_STATE, "swe_run: no state uploaded");
    // the chain starts from dt0 and zero time; kernels then carry it
    double init[2] = {dt0, 0.0};
    if (dt_clamp > 0.0 && dt_clamp < dt0) init[0] = dt_clamp;
    HIP_TRY(hipMemcpyAsync(&h->slots->dt, init, sizeof init, hipMemcpyHostToDevice, h->stream));

    auto one_step = [&](hipStream_t s, int from, int fwd) -> hipError_t {
        const int to = 1 - from;
        swe::StepArgs a = make_args(h, h->eta[from], h->hu[from], h->hv[from], h->b, h->eta[to],
                                    h->hu[to], h->hv[to], 0.0, fwd);
        a.device_dt = 1;
        a.dt_clamp = dt_clamp;
        return swe::launch_step(a, h->desc.solver, s);
    };

    int fwd = forward ? 1 : 0;
    const int64_t pairs = n_steps / 2;
    if (pairs > 0) {
        GraphCache &gc = h->graph;
        if (!gc.exec || gc.cur != h->cur || gc.forward != fwd || gc.dt_clamp != dt_clamp) {
            if (gc.exec) {
                (void)hipGraphExecDestroy(gc.exec);
                gc.exec = nullptr;
            }
            // capture on the library's own stream, replay on h->stream
            HIP_TRY(hipStreamSynchronize(h->stream));
            hipGraph_t graph;
            HIP_TRY(hipStreamBeginCapture(h->own_stream, hipStreamCaptureModeThreadLocal));
            hipError_t e1 = one_step(h->own_stream, h->cur, fwd);
            hipError_t e2 = e1 == hipSuccess ? one_step(h->own_stream, 1 - h->cur, 1 - fwd) : e1;
            hipError_t e3 = hipStreamEndCapture(h->own_stream, &graph);
            if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess)
                return fail(SWE_ERR_HIP, "swe_run: graph capture failed: %s",
                            hipGetErrorString(e1 != hipSuccess ? e1 : (e2 != hipSuccess ? e2 : e3)));
            hipError_t e4 = hipGraphInstantiate(&gc.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e4 != hipSuccess) return fail(SWE_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e4));
            gc.cur = h->cur;
            gc.forward = fwd;
            gc.dt_clamp = dt_clamp;
        }
        for (int64_t i = 0; i < pairs; ++i) HIP_TRY(hipGraphLaunch(gc.exec, h->stream));
    }
    if (n_steps % 2) {
        HIP_TRY(one_step(h->stream, h->cur, fwd));
        h->cur = 1 - h->cur;
    }
    double res[2];
    HIP_TRY(hipMemcpyAsync(res, &h->slots->dt, sizeof res, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (dt_next) *dt_next = res[0];
    if (time_added) *time_added = res[1];
    return SWE_OK;
}

int swe_set_stream(swe_handle *h, void *stream) {
    if (!h) return fail(SWE_ERR_ARG, "swe_set_stream: null handle");
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->stream = stream ? (hipStream_t)stream : h->own_stream;
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
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    swe::StepArgs a = make_args(h, eta, hu, hv, b, eta_out, hu_out, hv_out, dt, forward);
    a.refine = refine_dev;
    a.out_dtcfl = dt_cfl_dev;
    a.out_nref = n_refine_dev;
    HIP_TRY(swe::launch_step(a, h->desc.solver, s));
    return SWE_OK;
}

}  // extern "C"
