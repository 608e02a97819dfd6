// vah_batch.hip -- host side of the batch engine declared in include/vah_batch.h.
//
// Device memory:
//   arena        one allocation of budget_bytes; each resident event owns a
//                contiguous region (61 padded fields: 3 x 15 conserved,
//                e and p twice, u^mu three times), handed out first-fit
//   ev / dyn     per-slot event descriptors (host-written at admission) and
//                device-updated state (step counter, max e, counters, state)
//   tiles2/3     the 2+1D and 3+1D tile lists: (slot, tile) pairs, event by event
//   counts       [0] 2+1D tiles, [1] 3+1D tiles, [2] resident events
//   active       resident slots, for the advance kernel
// A step is: stage 1 (2+1D list, 3+1D list), stage 2 (both), advance.  Every
// launch reads its extent from `counts`, so a captured graph stays valid
// when events come and go; only a variant change or a larger tile list
// (reallocation) drops the captured graphs.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "../../include/vah_batch.h"
#include "../host/ic_internal.h"
#include "vah_batch_kernels.h"
#include "vah_glauber.h"
#include "vah_host_common.h"

using namespace vah;
using vah_host::fail;

namespace vah {
// one launch function per compiled variant (vah_batch_variant.hip)
void batch_stage_v16x15b2(const BatchLaunch &, bool, bool, bool);
void batch_stage_v32x7b2(const BatchLaunch &, bool, bool, bool);
void batch_stage_v32x7b3(const BatchLaunch &, bool, bool, bool);
void batch_stage_v16x15b3(const BatchLaunch &, bool, bool, bool);
}  // namespace vah

namespace {

constexpr double kHbarC = 0.197326938;  // GeV fm
constexpr int kFieldsPerEvent = 3 * NQ + 4 + 12;
constexpr size_t kAlign = 256;

struct Variant {
    vah_batch_variant info;
    BatchStageFn fn;
};

const Variant kVariants[] = {
    {{16, 15, 2, 3}, batch_stage_v16x15b2},
    {{16, 15, 2, 6}, batch_stage_v16x15b2},
    {{32, 7, 2, 3}, batch_stage_v32x7b2},
    {{32, 7, 2, 6}, batch_stage_v32x7b2},
    {{16, 15, 3, 3}, batch_stage_v16x15b3},
    {{32, 7, 3, 3}, batch_stage_v32x7b3},
};
constexpr int kNumVariants = (int)(sizeof kVariants / sizeof kVariants[0]);

struct HostEvent {
    bool used = false;
    bool stopped = false;  // seen stopped by a poll: its tiles leave the lists
    vah_batch_event desc;
    vah_host::Layout lay;
    size_t offset = 0, bytes = 0;
    BatchEvent dev;  // the descriptor as uploaded (pointers into the arena)
};

}  // namespace

struct vah_batch {
    vah_batch_options opt;
    int device;
    hipStream_t stream;
    bool own_stream;
    int num_cus;
    // EOS
    Eos eos;
    vah_host::HostEos heos;
    uint32_t *powf_fix;
    double *eos_dev;
    // arena and its free regions (offset -> size)
    char *arena;
    size_t arena_bytes, used_bytes;
    std::map<size_t, size_t> free_list;
    // event tables
    int max_events;
    std::vector<HostEvent> events;
    BatchEvent *ev_dev;
    BatchDyn *dyn_dev;
    BatchStageRec *rec_dev;  // per slot and stage, written at the start of each step
    int *active_dev;
    int *counts_dev;
    int2 *tiles_dev;   // tiles2 at [0, cap2), tiles3 at [cap2, cap2 + cap3)
    size_t cap2, cap3;
    BatchDyn *dyn_host;  // pinned read-back
    bool lists_dirty;
    int n2, n3, nactive;
    // kernels
    int variant;
    std::map<int, hipGraphExec_t> graphs;  // key: (has2d, has3d)
    // pinned staging buffers for admissions, reused round-robin once their
    // copies have completed
    struct Staging {
        char *host = nullptr;
        size_t cap = 0;
        hipEvent_t done = nullptr;
        bool pending = false;
    };
    Staging stage[4];
    int stage_next;
};

namespace {

size_t round_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

size_t event_bytes(int nx, int ny, int nz) {
    vah_host::Layout L(nx, ny, nz);
    return round_up(sizeof(double) * (size_t)L.fs * kFieldsPerEvent, kAlign);
}

// first-fit allocation in the arena; returns false when no region fits
bool arena_alloc(vah_batch *b, size_t bytes, size_t *off) {
    for (auto it = b->free_list.begin(); it != b->free_list.end(); ++it) {
        if (it->second >= bytes) {
            *off = it->first;
            size_t rest = it->second - bytes;
            size_t at = it->first + bytes;
            b->free_list.erase(it);
            if (rest) b->free_list[at] = rest;
            b->used_bytes += bytes;
            return true;
        }
    }
    return false;
}

void arena_free(vah_batch *b, size_t off, size_t bytes) {
    b->used_bytes -= bytes;
    auto it = b->free_list.emplace(off, bytes).first;
    // merge with the following and the preceding region
    auto nx = std::next(it);
    if (nx != b->free_list.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        b->free_list.erase(nx);
    }
    if (it != b->free_list.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            b->free_list.erase(it);
        }
    }
}

int tiles_of(const vah_batch *b, const HostEvent &h) {
    const vah_batch_variant &v = kVariants[b->variant].info;
    int tx = (h.desc.nx + v.tile_x - 1) / v.tile_x;
    int ty = (h.desc.ny + v.tile_y - 1) / v.tile_y;
    int kz = h.desc.nz > 1 ? std::min(v.kz, h.desc.nz) : 1;
    int zc = (h.desc.nz + kz - 1) / kz;
    return tx * ty * zc;
}

// the stage constants of one event under the current variant (the values
// vah_step's make_params would use, pointers left to batch_roles)
void fill_descriptor(const vah_batch *b, HostEvent &h) {
    const vah_batch_variant &v = kVariants[b->variant].info;
    const vah_batch_event &d = h.desc;
    StageParams &P = h.dev.base;
    memset(&P, 0, sizeof P);
    P.fs = h.lay.fs;
    P.nx = d.nx;
    P.ny = d.ny;
    P.nz = d.nz;
    P.ncx = h.lay.ncx;
    P.ncy = h.lay.ncy;
    P.sy = h.lay.ncx;
    P.sz = (long)h.lay.ncx * h.lay.ncy;
    P.tiles_x = (d.nx + v.tile_x - 1) / v.tile_x;
    P.tiles_y = (d.ny + v.tile_y - 1) / v.tile_y;
    P.kz = d.nz > 1 ? std::min(v.kz, d.nz) : 1;
    P.zchunks = (d.nz + P.kz - 1) / P.kz;
    P.nblocks = P.tiles_x * P.tiles_y * P.zchunks;
    P.dt = d.dt;
    P.dt_inv = 1.0 / d.dt;
    P.etabar = d.etabar;
    P.etabar_inv = 1.0 / d.etabar;
    P.inv_dx = 1.0 / d.dx;
    P.inv_dy = 1.0 / d.dy;
    P.inv_dz = 1.0 / d.dz;
    P.hx = 1 / d.dx / 2;
    P.hy = 1 / d.dy / 2;
    P.hz = 1 / d.dz / 2;
    P.dt_dx = P.dt * P.inv_dx;
    P.dt_dy = P.dt * P.inv_dy;
    P.dt_dz = P.dt * P.inv_dz;
    P.eos = b->eos;
    double *base = (double *)(b->arena + h.offset);
    const long fs = h.lay.fs;
    for (int k = 0; k < 3; ++k, base += NQ * fs) h.dev.q[k] = base;
    for (int k = 0; k < 2; ++k, base += fs) h.dev.e[k] = base;
    for (int k = 0; k < 2; ++k, base += fs) h.dev.p[k] = base;
    for (int k = 0; k < 3; ++k, base += 4 * fs) h.dev.u[k] = base;
    h.dev.t0 = d.t0;
    h.dev.dt = d.dt;
    h.dev.Tf = b->opt.freezeout_T_GeV > 0 ? b->opt.freezeout_T_GeV / kHbarC : 0.0;
    h.dev.max_steps = d.max_steps;
}

void drop_graphs(vah_batch *b) {
    for (auto &g : b->graphs) (void)hipGraphExecDestroy(g.second);
    b->graphs.clear();
}

// rebuild and upload the tile and slot lists (the stream is idle or the
// uploads are ordered after the work that reads the old lists)
int upload_lists(vah_batch *b) {
    std::vector<int2> t2, t3;
    std::vector<int> act, ev3;
    for (int s = 0; s < b->max_events; ++s) {
        const HostEvent &h = b->events[s];
        if (!h.used || h.stopped) continue;  // a stopped event's fields stay put without its tiles
        act.push_back(s);
        if (h.desc.nz > 1) {
            ev3.push_back(s);
            continue;
        }
        const int nt = tiles_of(b, h);
        for (int t = 0; t < nt; ++t) t2.push_back(make_int2(s, t));
    }
    // 3+1D tiles: the kernel gives XCD x the x-th eighth of the list, in list
    // order.  The list is therefore eight groups, group x holding the x-th
    // eighth of every eta chunk's tiles, chunk after chunk: the XCDs work
    // through each event's chunks together, so that the planes in flight form
    // one band, as in the single-lattice kernel's workgroup order.
    const vah_batch_variant &v = kVariants[b->variant].info;
    for (int x = 0; x < 8; ++x) {
        for (int s : ev3) {
            const HostEvent &h = b->events[s];
            const int tpl = ((h.desc.nx + v.tile_x - 1) / v.tile_x) * ((h.desc.ny + v.tile_y - 1) / v.tile_y);
            const int zc = tiles_of(b, h) / tpl;
            const int t0 = tpl * x / 8, t1 = tpl * (x + 1) / 8;
            for (int c = 0; c < zc; ++c)
                for (int t = t0; t < t1; ++t) t3.push_back(make_int2(s, c * tpl + t));
        }
    }
    if (t2.size() > b->cap2 || t3.size() > b->cap3) {
        HIP_TRY(hipStreamSynchronize(b->stream));
        drop_graphs(b);
        (void)hipFree(b->tiles_dev);
        b->tiles_dev = nullptr;
        b->cap2 = std::max(b->cap2, round_up(t2.size(), 4096));
        b->cap3 = std::max(b->cap3, round_up(t3.size(), 4096));
        if (hipMalloc(&b->tiles_dev, sizeof(int2) * (b->cap2 + b->cap3)) != hipSuccess)
            return fail(VAH_ERR_NOMEM, "vah_batch: tile list of %zu entries", b->cap2 + b->cap3);
    }
    if (!t2.empty())
        HIP_TRY(hipMemcpyAsync(b->tiles_dev, t2.data(), sizeof(int2) * t2.size(), hipMemcpyHostToDevice, b->stream));
    if (!t3.empty())
        HIP_TRY(hipMemcpyAsync(b->tiles_dev + b->cap2, t3.data(), sizeof(int2) * t3.size(), hipMemcpyHostToDevice,
                               b->stream));
    if (!act.empty())
        HIP_TRY(hipMemcpyAsync(b->active_dev, act.data(), sizeof(int) * act.size(), hipMemcpyHostToDevice, b->stream));
    int counts[3] = {(int)t2.size(), (int)t3.size(), (int)act.size()};
    HIP_TRY(hipMemcpyAsync(b->counts_dev, counts, sizeof counts, hipMemcpyHostToDevice, b->stream));
    // the host vectors die here: wait for the copies
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->n2 = counts[0];
    b->n3 = counts[1];
    b->nactive = counts[2];
    b->lists_dirty = false;
    return VAH_OK;
}

// persistent workgroups: the variant's occupancy, at most 2 per CU for 3+1D
// tiles (their face and eta buffers fill half a CU's LDS)
int grid_of(const vah_batch *b, bool d3) {
    const vah_batch_variant &v = kVariants[b->variant].info;
    int g = b->num_cus * (d3 ? std::min(v.blocks_per_cu, 2) : v.blocks_per_cu);
    return std::max(8, g / 8 * 8);
}

// enqueue one step of every live event
void enqueue_step(vah_batch *b, bool has2, bool has3) {
    const int blocks = (b->max_events + 63) / 64;
    hipLaunchKernelGGL(batch_prepare_kernel, dim3(blocks), dim3(64), 0, b->stream, b->ev_dev, b->dyn_dev,
                       b->active_dev, b->counts_dev + 2, b->rec_dev);
    BatchLaunch L;
    L.rec = b->rec_dev;
    L.tiles2 = b->tiles_dev;
    L.tiles3 = b->tiles_dev + b->cap2;
    L.counts = b->counts_dev;
    L.grid = grid_of(b, false);
    L.grid3 = grid_of(b, true);
    L.stream = b->stream;
    const bool fp32 = b->opt.precision == VAH_FP32_SPLIT;
    BatchStageFn fn = kVariants[b->variant].fn;
    for (int stage2 = 0; stage2 < 2; ++stage2) {
        if (has2) fn(L, false, stage2 != 0, fp32);
        if (has3) fn(L, true, stage2 != 0, fp32);
    }
    hipLaunchKernelGGL(batch_advance_kernel, dim3(blocks), dim3(64), 0, b->stream, b->ev_dev, b->dyn_dev,
                       b->active_dev, b->counts_dev + 2);
}

int graph_for(vah_batch *b, bool has2, bool has3, hipGraphExec_t *out) {
    const int key = (has2 ? 1 : 0) | (has3 ? 2 : 0);
    auto it = b->graphs.find(key);
    if (it != b->graphs.end()) {
        *out = it->second;
        return VAH_OK;
    }
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    for (int s = 0; s < b->opt.steps_per_graph; ++s) enqueue_step(b, has2, has3);
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (e != hipSuccess) return fail(VAH_ERR_HIP, "vah_batch: graph capture: %s", hipGetErrorString(e));
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(VAH_ERR_HIP, "vah_batch: graph instantiate: %s", hipGetErrorString(e));
    b->graphs[key] = x;
    *out = x;
    return VAH_OK;
}

bool slot_ok(const vah_batch *b, int slot) { return b && slot >= 0 && slot < b->max_events && b->events[slot].used; }

// Initial state of an admitted event from compact interior arrays
// [e | P_L? | u^tau u^x u^y u^eta? | pi (10)?] (each n values): the map of
// setConservedVariables (DynamicalVariables.cpp:197-287) -- without P_L the
// reference's 0.0154483 e, without u the rest frame, without pi zero -- into
// the padded level-0 buffers (q, e, p, u^n and u^{n-1} = u^n).
__global__ void init_event_kernel(const double *in, int has_pl, int has_u, int has_pi, int nx, int ny, int nz,
                                  long sy, long sz, long fs, double t, Eos eos, double *q, double *e, double *p,
                                  double *u0, double *u1) {
    const long n = (long)nx * ny * nz;
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const int ii = (int)(m % nx);
    const int jj = (int)((m / nx) % ny);
    const int kk = (int)(m / ((long)nx * ny));
    const long s = (long)plane_of(kk, nz > 1) * sz + (long)(jj + 2) * sy + (ii + 2);
    const double *col = in + n;
    const double eeq = in[m];
    double plv = 0.0154483 * eeq;
    if (has_pl) {
        plv = col[m];
        col += n;
    }
    double vt = 1.0, vx = 0.0, vy = 0.0, vn = 0.0;
    if (has_u) {
        vt = col[m];
        vx = col[n + m];
        vy = col[2 * n + m];
        vn = col[3 * n + m];
        col += 4 * n;
    }
    double pis[10];
#pragma unroll
    for (int f = 0; f < 10; ++f) pis[f] = has_pi ? col[f * n + m] : 0.0;
    const double pt = 0.5 * (eeq - plv);
    const double DP = plv - pt;
    const double F = 1 + (vx * vx + vy * vy);
    const double Ltt = DP * t * t * vn * vn / F;
    const double Ltn = DP * vt * vn / F;
    q[TTT * fs + s] = (eeq + pt) * vt * vt - pt + Ltt + pis[0];
    q[TTX * fs + s] = (eeq + pt) * vt * vx + pis[1];
    q[TTY * fs + s] = (eeq + pt) * vt * vy + pis[2];
    q[TTN * fs + s] = (eeq + pt) * vt * vn + Ltn + pis[3];
    q[PL * fs + s] = plv;
#pragma unroll
    for (int f = 0; f < 10; ++f) q[(PITT + f) * fs + s] = pis[f];
    e[s] = eeq;
    p[s] = eos_p(eos, eeq);
    const double uv[4] = {vt, vx, vy, vn};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        u0[k * fs + s] = uv[k];
        u1[k * fs + s] = uv[k];
    }
}

// a pinned staging buffer of at least `bytes`, free of pending copies
int staging(vah_batch *b, size_t bytes, vah_batch::Staging **out) {
    vah_batch::Staging &st = b->stage[b->stage_next];
    b->stage_next = (b->stage_next + 1) % 4;
    if (st.pending) {
        HIP_TRY(hipEventSynchronize(st.done));
        st.pending = false;
    }
    if (st.cap < bytes) {
        if (st.host) (void)hipHostFree(st.host);
        st.host = nullptr;
        st.cap = 0;
        // at least 8 MB: events of mixed sizes then rarely regrow a buffer
        // (each regrowth frees and allocates pinned memory, milliseconds apiece)
        const size_t cap = std::max(bytes + bytes / 4, (size_t)8 << 20);
        if (hipHostMalloc((void **)&st.host, cap, hipHostMallocDefault) != hipSuccess)
            return fail(VAH_ERR_NOMEM, "vah_batch_admit: %.1f MB of pinned staging", cap / 1e6);
        st.cap = cap;
    }
    if (!st.done) HIP_TRY(hipEventCreateWithFlags(&st.done, hipEventDisableTiming));
    *out = &st;
    return VAH_OK;
}

}  // namespace

extern "C" {

void vah_batch_default_options(vah_batch_options *opt) {
    if (!opt) return;
    memset(opt, 0, sizeof *opt);
    opt->eos = VAH_EOS_CONFORMAL;
    opt->precision = VAH_FP64;
    opt->budget_bytes = (size_t)8 << 30;
    opt->max_events = 1024;
    opt->freezeout_T_GeV = 0.150;
    opt->steps_per_graph = 16;
    opt->variant = -1;
    opt->device = 0;
    opt->stream = nullptr;
}

size_t vah_batch_event_bytes(int nx, int ny, int nz) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return 0;
    return event_bytes(nx, ny, nz);
}

int vah_batch_variant_count(void) { return kNumVariants; }

int vah_batch_variant_info(int index, vah_batch_variant *out) {
    if (!out || index < 0 || index >= kNumVariants) return fail(VAH_ERR_ARG, "vah_batch_variant_info: bad index");
    *out = kVariants[index].info;
    return VAH_OK;
}

void vah_batch_destroy(vah_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graphs(b);
    (void)hipFree(b->arena);
    (void)hipFree(b->ev_dev);
    (void)hipFree(b->dyn_dev);
    (void)hipFree(b->rec_dev);
    (void)hipFree(b->active_dev);
    (void)hipFree(b->counts_dev);
    (void)hipFree(b->tiles_dev);
    (void)hipFree(b->powf_fix);
    (void)hipFree(b->eos_dev);
    if (b->dyn_host) (void)hipHostFree(b->dyn_host);
    for (auto &st : b->stage) {
        if (st.host) (void)hipHostFree(st.host);
        if (st.done) (void)hipEventDestroy(st.done);
    }
    if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int vah_batch_create(const vah_batch_options *opt_in, vah_batch **out) {
    if (!out) return fail(VAH_ERR_ARG, "vah_batch_create: out is NULL");
    *out = nullptr;
    vah_batch_options opt;
    if (opt_in)
        opt = *opt_in;
    else
        vah_batch_default_options(&opt);
    if (opt.precision != VAH_FP64 && opt.precision != VAH_FP32_SPLIT)
        return fail(VAH_ERR_ARG, "vah_batch_create: unknown precision %d", opt.precision);
    if (opt.max_events <= 0 || opt.max_events > (1 << 20))
        return fail(VAH_ERR_ARG, "vah_batch_create: max_events must be in [1, 2^20]");
    if (opt.budget_bytes < kAlign) return fail(VAH_ERR_ARG, "vah_batch_create: budget too small");
    if (opt.steps_per_graph < 0 || opt.steps_per_graph > 4096)
        return fail(VAH_ERR_ARG, "vah_batch_create: steps_per_graph must be in [0, 4096]");
    if (opt.variant < -1 || opt.variant >= kNumVariants)
        return fail(VAH_ERR_ARG, "vah_batch_create: unknown variant %d", opt.variant);
    if (opt.eos == VAH_EOS_TABLE && (opt.table.n < 2 || !opt.table.p || !opt.table.T || !(opt.table.h > 0)))
        return fail(VAH_ERR_ARG, "vah_batch_create: EOS table needs n >= 2 nodes, h > 0 and p, T arrays");
    if (opt.eos != VAH_EOS_CONFORMAL && opt.eos != VAH_EOS_TABLE && opt.eos != VAH_EOS_WB)
        return fail(VAH_ERR_ARG, "vah_batch_create: unknown EOS kind %d", opt.eos);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= opt.device || opt.device < 0)
        return fail(VAH_ERR_DEVICE, "vah_batch_create: HIP device %d not available (%d found)", opt.device, ndev);
    HIP_TRY(hipSetDevice(opt.device));

    vah_batch *b = new vah_batch();
    b->opt = opt;
    b->device = opt.device;
    b->max_events = opt.max_events;
    b->events.resize(opt.max_events);
    b->variant = opt.variant < 0 ? 0 : opt.variant;
    b->lists_dirty = true;
    auto cleanup = [&](int code) {
        vah_batch_destroy(b);
        return code;
    };
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, opt.device) != hipSuccess)
        return cleanup(fail(VAH_ERR_HIP, "vah_batch_create: device properties"));
    b->num_cus = prop.multiProcessorCount;
    if (opt.stream) {
        b->stream = (hipStream_t)opt.stream;
        b->own_stream = false;
    } else {
        if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
            return cleanup(fail(VAH_ERR_HIP, "vah_batch_create: stream creation failed"));
        b->own_stream = true;
    }
    b->arena_bytes = opt.budget_bytes / kAlign * kAlign;
    if (hipMalloc(&b->arena, b->arena_bytes) != hipSuccess)
        return cleanup(fail(VAH_ERR_NOMEM, "vah_batch_create: cannot allocate the %.2f GB arena", b->arena_bytes / 1e9));
    b->free_list[0] = b->arena_bytes;
    const size_t ne = (size_t)opt.max_events;
    if (hipMalloc(&b->ev_dev, sizeof(BatchEvent) * ne) != hipSuccess ||
        hipMalloc(&b->dyn_dev, sizeof(BatchDyn) * ne) != hipSuccess ||
        hipMalloc(&b->rec_dev, sizeof(BatchStageRec) * 2 * ne) != hipSuccess ||
        hipMalloc(&b->active_dev, sizeof(int) * ne) != hipSuccess ||
        hipMalloc(&b->counts_dev, sizeof(int) * 4) != hipSuccess ||
        hipHostMalloc(&b->dyn_host, sizeof(BatchDyn) * ne) != hipSuccess)
        return cleanup(fail(VAH_ERR_NOMEM, "vah_batch_create: event tables"));
    if (hipMemsetAsync(b->dyn_dev, 0, sizeof(BatchDyn) * ne, b->stream) != hipSuccess ||
        hipMemsetAsync(b->rec_dev, 0, sizeof(BatchStageRec) * 2 * ne, b->stream) != hipSuccess ||
        hipMemsetAsync(b->counts_dev, 0, sizeof(int) * 4, b->stream) != hipSuccess)
        return cleanup(fail(VAH_ERR_HIP, "vah_batch_create: table clear"));
    b->cap2 = b->cap3 = 4096;
    if (hipMalloc(&b->tiles_dev, sizeof(int2) * (b->cap2 + b->cap3)) != hipSuccess)
        return cleanup(fail(VAH_ERR_NOMEM, "vah_batch_create: tile lists"));
    std::string err;
    int rc = vah_host::make_device_eos(opt.eos, opt.table, b->stream, &b->eos, &b->powf_fix, &b->eos_dev, &b->heos,
                                       &err);
    if (rc != VAH_OK) return cleanup(fail(rc, "vah_batch_create: %s", err.c_str()));
    b->opt.table.p = nullptr;
    b->opt.table.T = nullptr;
    if (hipStreamSynchronize(b->stream) != hipSuccess) return cleanup(fail(VAH_ERR_HIP, "vah_batch_create: sync"));
    *out = b;
    return VAH_OK;
}

}  // extern "C"

namespace {

// Checks an event's description, takes a free slot and an arena region for it
// and fills its descriptor; the slot is not marked used yet.
int reserve_event(vah_batch *b, const vah_batch_event *d, const char *who, int *slot_out) {
    if (d->nx <= 0 || d->ny <= 0 || d->nz <= 0 || !(d->dx > 0) || !(d->dy > 0) || !(d->dz > 0) || !(d->dt > 0))
        return fail(VAH_ERR_ARG, "%s: lattice sizes and spacings must be positive", who);
    if (!(d->t0 > 0) || !(d->etabar > 0)) return fail(VAH_ERR_ARG, "%s: need t0 > 0 and eta/s > 0", who);
    int s = -1;
    for (int k = 0; k < b->max_events; ++k)
        if (!b->events[k].used) {
            s = k;
            break;
        }
    if (s < 0) return fail(VAH_ERR_FULL, "%s: all %d event slots are taken", who, b->max_events);
    if ((double)vah_host::Layout(d->nx, d->ny, d->nz).fs * NQ * 8 >= 4294967296.0)
        return fail(VAH_ERR_ARG, "%s: a 15-field set of this lattice exceeds 4 GB", who);
    const size_t bytes = event_bytes(d->nx, d->ny, d->nz);
    size_t off = 0;
    if (!arena_alloc(b, bytes, &off))
        return fail(VAH_ERR_BUDGET, "%s: %.1f MB do not fit (%.1f of %.1f MB in use)", who, bytes / 1e6,
                    b->used_bytes / 1e6, b->arena_bytes / 1e6);
    HostEvent &h = b->events[s];
    h.desc = *d;
    h.lay = vah_host::Layout(d->nx, d->ny, d->nz);
    h.offset = off;
    h.bytes = bytes;
    fill_descriptor(b, h);
    *slot_out = s;
    return VAH_OK;
}

void release_event(vah_batch *b, int s) {
    HostEvent &h = b->events[s];
    arena_free(b, h.offset, h.bytes);
    h.used = false;
}

// Level-0 state of event s from compact inputs already on the device at
// dev_in (init_event_kernel's layout), its ghosts, and the upload of its
// descriptor and dynamic record from the staging buffer; marks the slot used.
int finish_admission(vah_batch *b, int s, const double *dev_in, int has_pl, int has_u, int has_pi,
                     vah_batch::Staging *st, size_t desc_off, size_t dyn_off, const char *who) {
    HostEvent &h = b->events[s];
    const vah_batch_event *d = &h.desc;
    const long n = (long)d->nx * d->ny * d->nz;
    BatchDyn dyn;
    memset(&dyn, 0, sizeof dyn);
    dyn.live = 1;
    dyn.state = EV_LIVE;
    memcpy(st->host + desc_off, &h.dev, sizeof(BatchEvent));
    memcpy(st->host + dyn_off, &dyn, sizeof(BatchDyn));
    const int blocks = (int)((n + 255) / 256);
    hipLaunchKernelGGL(init_event_kernel, dim3(blocks), dim3(256), 0, b->stream, dev_in, has_pl, has_u, has_pi,
                       d->nx, d->ny, d->nz, (long)h.lay.ncx, (long)h.lay.ncx * h.lay.ncy, h.lay.fs, d->t0, b->eos,
                       h.dev.q[0], h.dev.e[0], h.dev.p[0], h.dev.u[0], h.dev.u[1]);
    hipLaunchKernelGGL(ghost_fill_kernel, dim3(blocks), dim3(256), 0, b->stream, h.dev.q[0], h.dev.e[0], h.dev.p[0],
                       h.dev.u[0], h.lay.fs, d->nx, d->ny, d->nz, (long)h.lay.ncx, (long)h.lay.ncx * h.lay.ncy,
                       d->nz > 1 ? 1 : 0);
    if (hipGetLastError() != hipSuccess) {
        release_event(b, s);
        return fail(VAH_ERR_HIP, "%s: initial-state launch", who);
    }
    if (hipMemcpyAsync(b->ev_dev + s, st->host + desc_off, sizeof(BatchEvent), hipMemcpyHostToDevice, b->stream) !=
            hipSuccess ||
        hipMemcpyAsync(b->dyn_dev + s, st->host + dyn_off, sizeof(BatchDyn), hipMemcpyHostToDevice, b->stream) !=
            hipSuccess ||
        hipEventRecord(st->done, b->stream) != hipSuccess) {
        release_event(b, s);
        return fail(VAH_ERR_HIP, "%s: descriptor upload", who);
    }
    st->pending = true;
    h.used = true;
    h.stopped = false;
    b->lists_dirty = true;
    return VAH_OK;
}

}  // namespace

extern "C" {

int vah_batch_admit(vah_batch *b, const vah_batch_event *d, const double *e, const double *pl, const double *ut,
                    const double *ux, const double *uy, const double *un, const double *pi, int *slot) {
    if (!b || !d || !e || !slot) return fail(VAH_ERR_ARG, "vah_batch_admit: NULL argument");
    *slot = -1;
    const int nu = (ut != nullptr) + (ux != nullptr) + (uy != nullptr) + (un != nullptr);
    if (nu != 0 && nu != 4) return fail(VAH_ERR_ARG, "vah_batch_admit: give all four u^mu arrays or none");
    HIP_TRY(hipSetDevice(b->device));
    const long n = d->nx > 0 && d->ny > 0 && d->nz > 0 ? (long)d->nx * d->ny * d->nz : 0;
    const int nin = 1 + (pl ? 1 : 0) + (nu ? 4 : 0) + (pi ? 10 : 0);
    const size_t in_bytes = sizeof(double) * (size_t)nin * n;
    const size_t desc_off = (in_bytes + 255) / 256 * 256;
    const size_t dyn_off = desc_off + (sizeof(BatchEvent) + 255) / 256 * 256;
    int rc, s = -1;
    if ((rc = reserve_event(b, d, "vah_batch_admit", &s)) != VAH_OK) return rc;
    vah_batch::Staging *st = nullptr;
    if ((rc = staging(b, dyn_off + sizeof(BatchDyn), &st)) != VAH_OK) {
        release_event(b, s);
        return rc;
    }
    HostEvent &h = b->events[s];
    // compact inputs into pinned memory, then everything else on the device:
    // the inputs land in the event's own q* and q^{n+1} sets (free until its
    // first step), a kernel maps them to the padded level-0 state
    double *col = (double *)st->host;
    auto put = [&](const double *src) {
        memcpy(col, src, sizeof(double) * n);
        col += n;
    };
    put(e);
    if (pl) put(pl);
    if (nu) {
        put(ut);
        put(ux);
        put(uy);
        put(un);
    }
    if (pi) {
        memcpy(col, pi, sizeof(double) * 10 * n);
        col += 10 * n;
    }
    double *dev_in = h.dev.q[1];  // 2 x 15 fields of fs >= nin * n values
    if (hipMemcpyAsync(dev_in, st->host, in_bytes, hipMemcpyHostToDevice, b->stream) != hipSuccess) {
        release_event(b, s);
        return fail(VAH_ERR_HIP, "vah_batch_admit: input upload");
    }
    if ((rc = finish_admission(b, s, dev_in, pl ? 1 : 0, nu ? 1 : 0, pi ? 1 : 0, st, desc_off, dyn_off,
                               "vah_batch_admit")) != VAH_OK)
        return rc;
    *slot = s;
    return VAH_OK;
}

int vah_batch_admit_glauber(vah_batch *b, const vah_batch_event *d, const vah_glauber *g, double pl_ratio,
                            int *slot) {
    static const char *who = "vah_batch_admit_glauber";
    if (!b || !d || !g || !slot) return fail(VAH_ERR_ARG, "%s: NULL argument", who);
    *slot = -1;
    if (g->A < 1 || g->A > 1000 || !(g->R > 0) || !(g->a > 0) || !(g->nucleon_width > 0) || !(g->sigma_nn_mb >= 0))
        return fail(VAH_ERR_ARG, "%s: need 1 <= A <= 1000, R, a, nucleon width > 0, sigma_NN >= 0", who);
    HIP_TRY(hipSetDevice(b->device));
    // the host part: 2A nucleon positions from the seed (vah_ic_mc_glauber's draws)
    std::vector<double> xa, ya, xb, yb;
    if (vah_host::mc_glauber_nucleons(*g, xa, ya, xb, yb) != VAH_OK) return fail(VAH_ERR_ARG, "%s: geometry", who);
    const int A = g->A;
    if (d->nx <= 0 || d->ny <= 0 || d->nz <= 0) return fail(VAH_ERR_ARG, "%s: lattice sizes must be positive", who);
    const long n = (long)d->nx * d->ny * d->nz, nT = (long)d->nx * d->ny;
    // staging: positions (4A), eta factors (nz), descriptor, dynamic record
    const size_t pos_bytes = sizeof(double) * (4 * (size_t)A + d->nz);
    const size_t desc_off = (pos_bytes + 255) / 256 * 256;
    const size_t dyn_off = desc_off + (sizeof(BatchEvent) + 255) / 256 * 256;
    int rc, s = -1;
    if ((rc = reserve_event(b, d, who, &s)) != VAH_OK) return rc;
    vah_batch::Staging *st = nullptr;
    if ((rc = staging(b, dyn_off + sizeof(BatchDyn), &st)) != VAH_OK) {
        release_event(b, s);
        return rc;
    }
    HostEvent &h = b->events[s];
    // device scratch inside the event's q* and q^{n+1} sets (2 x 15 fields, free
    // until the first step): [e | P_L] (2n), e_T (nT), max, positions, eta
    // factors, wounded flags
    const size_t align = 32;  // doubles
    const size_t o_eT = ((size_t)2 * n + align - 1) / align * align;
    const size_t o_max = (o_eT + nT + align - 1) / align * align;
    const size_t o_pos = o_max + align;
    const size_t o_flag = (o_pos + 4 * (size_t)A + d->nz + align - 1) / align * align;
    const size_t need = o_flag + (2 * (size_t)A + 1) / 2 + 1;
    if (need > (size_t)2 * NQ * h.lay.fs) {
        release_event(b, s);
        return fail(VAH_ERR_ARG, "%s: lattice too small for the generator's scratch (%zu of %ld values)", who, need,
                    2L * NQ * h.lay.fs);
    }
    double *base = h.dev.q[1];
    double *dev_in = base, *eT = base + o_eT, *pos = base + o_pos, *fl = pos + 4 * A;
    unsigned long long *emax = (unsigned long long *)(base + o_max);
    int *wounded = (int *)(base + o_flag);
    double *hp = (double *)st->host;
    memcpy(hp, xa.data(), sizeof(double) * A);
    memcpy(hp + A, ya.data(), sizeof(double) * A);
    memcpy(hp + 2 * A, xb.data(), sizeof(double) * A);
    memcpy(hp + 3 * A, yb.data(), sizeof(double) * A);
    for (int k = 0; k < d->nz; ++k) {
        const double eta = (k - (d->nz - 1) / 2.) * d->dz;
        hp[4 * A + k] = d->nz > 1 ? vah_ic_eta_profile(eta, g->eta_flat, g->eta_sigma) : 1.0;
    }
    if (hipMemcpyAsync(pos, st->host, pos_bytes, hipMemcpyHostToDevice, b->stream) != hipSuccess) {
        release_event(b, s);
        return fail(VAH_ERR_HIP, "%s: position upload", who);
    }
    const double d2 = 0.1 * g->sigma_nn_mb / 3.14159265358979323846;
    const double w = g->nucleon_width;
    hipLaunchKernelGGL(glauber_wounded_kernel, dim3((2 * A + 255) / 256), dim3(256), 0, b->stream, pos, A, d2, wounded,
                       emax);
    hipLaunchKernelGGL(glauber_deposit_kernel, dim3((unsigned)((nT + 255) / 256)), dim3(256), 0, b->stream, pos,
                       wounded, A, d->nx, d->ny, d->dx, d->dy, 1 / (2 * w * w), 64 * w * w, eT, emax);
    hipLaunchKernelGGL(glauber_extrude_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, b->stream, eT, emax,
                       fl, nT, n, g->e_peak, g->e_floor, pl_ratio, dev_in);
    if (hipGetLastError() != hipSuccess) {
        release_event(b, s);
        return fail(VAH_ERR_HIP, "%s: generator launch", who);
    }
    if ((rc = finish_admission(b, s, dev_in, pl_ratio > 0 ? 1 : 0, 0, 0, st, desc_off, dyn_off, who)) != VAH_OK)
        return rc;
    *slot = s;
    return VAH_OK;
}

int vah_batch_step(vah_batch *b, int nsteps) {
    if (!b) return fail(VAH_ERR_ARG, "vah_batch_step: NULL batch");
    if (nsteps < 0) return fail(VAH_ERR_ARG, "vah_batch_step: negative step count");
    HIP_TRY(hipSetDevice(b->device));
    int rc;
    if (b->lists_dirty && (rc = upload_lists(b)) != VAH_OK) return rc;
    if (b->nactive == 0 || nsteps == 0) return VAH_OK;
    const bool has2 = b->n2 > 0, has3 = b->n3 > 0;
    int left = nsteps;
    if (b->opt.steps_per_graph > 0) {
        while (left >= b->opt.steps_per_graph) {
            hipGraphExec_t g;
            if ((rc = graph_for(b, has2, has3, &g)) != VAH_OK) return rc;
            HIP_TRY(hipGraphLaunch(g, b->stream));
            left -= b->opt.steps_per_graph;
        }
    }
    for (; left > 0; --left) enqueue_step(b, has2, has3);
    HIP_TRY(hipGetLastError());
    return VAH_OK;
}

int vah_batch_synchronize(vah_batch *b) {
    if (!b) return fail(VAH_ERR_ARG, "vah_batch_synchronize: NULL batch");
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return VAH_OK;
}

int vah_batch_poll(vah_batch *b, vah_batch_status *out, int cap, int *n) {
    if (!b || !n || (cap > 0 && !out)) return fail(VAH_ERR_ARG, "vah_batch_poll: NULL argument");
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipMemcpyAsync(b->dyn_host, b->dyn_dev, sizeof(BatchDyn) * b->max_events, hipMemcpyDeviceToHost,
                           b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    int k = 0;
    for (int s = 0; s < b->max_events; ++s) {
        HostEvent &h = b->events[s];
        if (!h.used) continue;
        if (b->dyn_host[s].state != EV_LIVE && !h.stopped) {
            h.stopped = true;  // drop its tiles from the lists at the next step
            b->lists_dirty = true;
        }
        if (k < cap) {
            const BatchDyn &d = b->dyn_host[s];
            vah_batch_status &o = out[k];
            o.slot = s;
            o.state = d.state;
            o.tag = h.desc.tag;
            o.nstep = d.nstep;
            o.tau = h.desc.t0 + (double)d.nstep * h.desc.dt;
            o.max_T = d.maxT;
            o.recovery_fallbacks = d.counters[1];
            o.nan_cells = d.counters[0];
        }
        ++k;
    }
    *n = k;
    return VAH_OK;
}

int vah_batch_get_state(vah_batch *b, int slot, double *q, double *e, double *p, double *u, double *up) {
    if (!slot_ok(b, slot)) return fail(VAH_ERR_ARG, "vah_batch_get_state: no event in slot %d", slot);
    HIP_TRY(hipSetDevice(b->device));
    BatchDyn d;
    HIP_TRY(hipMemcpyAsync(&d, b->dyn_dev + slot, sizeof d, hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    const HostEvent &h = b->events[slot];
    const long long n = d.nstep;
    const int a = (n & 1) ? 2 : 0, uc = (int)(n & 1);
    std::vector<double> buf;
    auto get = [&](const double *src, int nf, double *dst) -> int {
        if (!dst) return VAH_OK;
        buf.resize((size_t)nf * h.lay.fs);
        HIP_TRY(hipMemcpyAsync(buf.data(), src, sizeof(double) * nf * h.lay.fs, hipMemcpyDeviceToHost, b->stream));
        HIP_TRY(hipStreamSynchronize(b->stream));
        h.lay.unpad(nf, buf.data(), dst);
        return VAH_OK;
    };
    int rc;
    if ((rc = get(h.dev.q[a], NQ, q)) != VAH_OK) return rc;
    if ((rc = get(h.dev.e[0], 1, e)) != VAH_OK) return rc;
    if ((rc = get(h.dev.p[0], 1, p)) != VAH_OK) return rc;
    if ((rc = get(h.dev.u[uc], 4, u)) != VAH_OK) return rc;
    if ((rc = get(h.dev.u[1 - uc], 4, up)) != VAH_OK) return rc;
    return VAH_OK;
}

int vah_batch_retire(vah_batch *b, int slot) {
    if (!slot_ok(b, slot)) return fail(VAH_ERR_ARG, "vah_batch_retire: no event in slot %d", slot);
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipStreamSynchronize(b->stream));
    HostEvent &h = b->events[slot];
    BatchDyn dyn;
    memset(&dyn, 0, sizeof dyn);
    HIP_TRY(hipMemcpyAsync(b->dyn_dev + slot, &dyn, sizeof dyn, hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    arena_free(b, h.offset, h.bytes);
    h.used = false;
    h.stopped = false;
    b->lists_dirty = true;
    return VAH_OK;
}

int vah_batch_memory(const vah_batch *b, size_t *used, size_t *budget) {
    if (!b) return fail(VAH_ERR_ARG, "vah_batch_memory: NULL batch");
    if (used) *used = b->used_bytes;
    if (budget) *budget = b->arena_bytes;
    return VAH_OK;
}

int vah_batch_set_variant(vah_batch *b, int index) {
    if (!b || index < 0 || index >= kNumVariants) return fail(VAH_ERR_ARG, "vah_batch_set_variant: bad variant");
    if (index == b->variant) return VAH_OK;
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipStreamSynchronize(b->stream));
    drop_graphs(b);
    b->variant = index;
    for (int s = 0; s < b->max_events; ++s) {
        HostEvent &h = b->events[s];
        if (!h.used) continue;
        fill_descriptor(b, h);
        HIP_TRY(hipMemcpyAsync(b->ev_dev + s, &h.dev, sizeof(BatchEvent), hipMemcpyHostToDevice, b->stream));
        HIP_TRY(hipStreamSynchronize(b->stream));
    }
    b->lists_dirty = true;
    return VAH_OK;
}

int vah_batch_get_variant(const vah_batch *b) { return b ? b->variant : -1; }

int vah_batch_autotune(vah_batch *b, int steps_per_trial, int *best, double *ms_per_step) {
    if (!b || !best || steps_per_trial <= 0) return fail(VAH_ERR_ARG, "vah_batch_autotune: bad argument");
    HIP_TRY(hipSetDevice(b->device));
    const int keep_graph = b->opt.steps_per_graph;
    b->opt.steps_per_graph = 0;  // time the kernels, not graph instantiation
    hipEvent_t t0, t1;
    HIP_TRY(hipEventCreate(&t0));
    HIP_TRY(hipEventCreate(&t1));
    int rc = VAH_OK, arg = -1;
    double best_ms = 0;
    for (int v = 0; v < kNumVariants && rc == VAH_OK; ++v) {
        if ((rc = vah_batch_set_variant(b, v)) != VAH_OK) break;
        if ((rc = vah_batch_step(b, 1)) != VAH_OK) break;  // list upload and first-launch costs
        if (hipEventRecord(t0, b->stream) != hipSuccess) { rc = fail(VAH_ERR_HIP, "autotune: event"); break; }
        if ((rc = vah_batch_step(b, steps_per_trial)) != VAH_OK) break;
        if (hipEventRecord(t1, b->stream) != hipSuccess || hipEventSynchronize(t1) != hipSuccess) {
            rc = fail(VAH_ERR_HIP, "autotune: event");
            break;
        }
        float ms = 0;
        (void)hipEventElapsedTime(&ms, t0, t1);
        const double per = ms / steps_per_trial;
        if (ms_per_step) ms_per_step[v] = per;
        if (arg < 0 || per < best_ms) {
            arg = v;
            best_ms = per;
        }
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    b->opt.steps_per_graph = keep_graph;
    if (rc != VAH_OK) return rc;
    *best = arg;
    return vah_batch_set_variant(b, arg);
}

}  // extern "C"
