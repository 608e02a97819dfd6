// vah_capi.hip -- host side of the C ABI declared in include/vah.h.
//
// One vah_ctx owns the device state of one lattice and a HIP stream.  A step
// is two launches of the fused stage kernel (vah_stage.h); buffer roles
// rotate on the host instead of copying:
//   conserved  q^n, q*, q^{n+1}      three 15-field slabs, q^n <-> q^{n+1} per step
//   velocity   u^n, u^{n-1}, u*      three 4-field slabs, (u^n, u^{n-1}) <- (old u^{n-1} slab, u^n)
//   e, p       two buffers each, stage 1 writes the other one, stage 2 writes back
// which is the data flow of rungeKutta2
// (old_cpu-vah_code/src/FullyDiscreteKurganovTadmorScheme.cpp:738-792) with the
// in-place e/p updates of the reference split into read and write buffers so
// that a cell's neighbours are never overwritten while being read.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/vah.h"
#include "vah_host_common.h"
#include "vah_stage.h"

using namespace vah;

namespace {
thread_local std::string g_err;
}  // namespace

namespace vah_host {

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

}  // namespace vah_host

using vah_host::fail;

// ---- glibc-compatible powf(x, 0.25f) correction bitmap -------------------
// Bit `off` is set when the host libm's powf differs from the correctly
// rounded fourth root for x with bit pattern (124 << 23) + off.  The pattern
// repeats every 4 binades (see conformal_T in vah_device.h).
namespace vah_host {

const std::vector<uint32_t> &powf_fix_table() {
    static std::once_flag once;
    static std::vector<uint32_t> fix;
    std::call_once(once, [] {
        const uint32_t nbits = 1u << 25;
        fix.assign(nbits / 32, 0u);
        unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> th;
        const uint32_t words = nbits / 32, per = (words + nth - 1) / nth;
        for (unsigned t = 0; t < nth; ++t) {
            th.emplace_back([t, per, words] {
                uint32_t w0 = t * per, w1 = std::min(words, w0 + per);
                for (uint32_t w = w0; w < w1; ++w) {
                    uint32_t word = 0;
                    for (uint32_t bit = 0; bit < 32; ++bit) {
                        uint32_t b = (124u << 23) + w * 32 + bit;
                        float x;
                        memcpy(&x, &b, 4);
                        float g = powf(x, 0.25f);
                        float r = (float)sqrt(sqrt((double)x));
                        if (g != r) word |= 1u << bit;
                    }
                    fix[w] = word;
                }
            });
        }
        for (auto &x : th) x.join();
    });
    return fix;
}

int make_device_eos(int kind, const vah_eos_table &T, hipStream_t stream, Eos *out, uint32_t **fix_dev,
                    double **tab_dev, HostEos *host, std::string *err) {
    memset(out, 0, sizeof *out);
    *fix_dev = nullptr;
    *tab_dev = nullptr;
    host->kind = kind;
    out->kind = kind;
    if (kind == VAH_EOS_WB) return VAH_OK;  // closed-form fits, nothing to upload
    if (kind == VAH_EOS_CONFORMAL) {
        const std::vector<uint32_t> &fix = powf_fix_table();
        if (hipMalloc(fix_dev, fix.size() * sizeof(uint32_t)) != hipSuccess) {
            *err = "powf table allocation";
            return VAH_ERR_NOMEM;
        }
        if (hipMemcpyAsync(*fix_dev, fix.data(), fix.size() * sizeof(uint32_t), hipMemcpyHostToDevice, stream) !=
                hipSuccess ||
            hipStreamSynchronize(stream) != hipSuccess) {
            *err = "powf table upload";
            return VAH_ERR_HIP;
        }
        out->powf_fix = *fix_dev;
        return VAH_OK;
    }
    if (hipMalloc(tab_dev, sizeof(double) * 2 * T.n) != hipSuccess) {
        *err = "EOS table allocation";
        return VAH_ERR_NOMEM;
    }
    std::vector<double> tab(2 * (size_t)T.n);
    memcpy(tab.data(), T.p, sizeof(double) * T.n);
    memcpy(tab.data() + T.n, T.T, sizeof(double) * T.n);
    if (hipMemcpy(*tab_dev, tab.data(), sizeof(double) * 2 * T.n, hipMemcpyHostToDevice) != hipSuccess) {
        *err = "EOS table upload";
        return VAH_ERR_HIP;
    }
    out->n = T.n;
    out->lne0 = T.lne0;
    out->inv_h = 1.0 / T.h;
    out->e_first = exp(T.lne0);
    out->e_last = exp(T.lne0 + (T.n - 1) / (1.0 / T.h));
    out->p = *tab_dev;
    out->T = *tab_dev + T.n;
    host->n = T.n;
    host->lne0 = T.lne0;
    host->h = T.h;
    host->p.assign(T.p, T.p + T.n);
    return VAH_OK;
}

}  // namespace vah_host

struct vah_ctx {
    vah_lattice lat;
    vah_hydro hydro;
    vah_options opt;
    int ncx, ncy, ncz;
    long ncell, fs;
    int device;
    hipStream_t stream;
    bool own_stream;
    double *slab;
    double *qset[3];  // conserved slabs
    double *useq[3];  // velocity slabs
    double *ebuf[2], *pbuf[2];
    int qi, qs, qo;        // q^n, q*, q^{n+1}
    int uc, up, us;        // u^n, u^{n-1}, u*
    int ecur;              // e/p buffer holding the current level
    unsigned int *counters;  // device, [0] NaN, [1] fallbacks
    unsigned int counters_seen[2];
    uint32_t *powf_fix;
    double *eos_dev;       // p then T
    Eos eos;
    vah_host::HostEos heos;
    vah_host::Layout lay;
    long long nstep;
    bool have_state;
    bool stepped;          // a step has completed since the last vah_set_state
    // launch geometry
    int tiles_x, tiles_y, zchunks, kz, nblocks, grid;
    // profiling
    bool prof;
    std::vector<hipEvent_t> ev;  // pairs per launch
    std::vector<int> ev_kind;
    double *red_buf;       // per-block partial reductions
    int red_blocks;
};

namespace {

bool lattice_ok(const vah_lattice *l) {
    return l && l->nx > 0 && l->ny > 0 && l->nz > 0 && l->dx > 0 && l->dy > 0 && l->dz > 0 && l->dt > 0;
}

// Launch geometry of one lattice: 16 x 15 tiles marching kz eta planes.
struct Geometry {
    int tiles_x, tiles_y, zchunks, kz, nblocks, grid;
};

Geometry make_geometry(const vah_lattice &lat, int eta_chunk) {
    Geometry g;
    g.tiles_x = (lat.nx + kTileX - 1) / kTileX;
    g.tiles_y = (lat.ny + kTileY - 1) / kTileY;
    // eta planes per workgroup: 3 by default (fastest in a sweep at 201x201x101)
    const int kz_default = 3;
    g.kz = lat.nz > 1 ? std::min(eta_chunk > 0 ? eta_chunk : kz_default, lat.nz) : 1;
    g.zchunks = (lat.nz + g.kz - 1) / g.kz;
    g.nblocks = g.tiles_x * g.tiles_y * g.zchunks;
    g.grid = g.zchunks * 8 * ((g.tiles_x * g.tiles_y + 7) / 8);  // tile_of_block's grid
    return g;
}

// Everything of a stage's parameters except its buffers and times.
StageParams base_params(const vah_lattice &lat, const vah_hydro &hydro, const vah_host::Layout &lay,
                        const Geometry &g, const Eos &eos, unsigned int *counters) {
    StageParams P;
    P.fs = lay.fs;
    P.nx = lat.nx;
    P.ny = lat.ny;
    P.nz = lat.nz;
    P.ncx = lay.ncx;
    P.ncy = lay.ncy;
    P.sy = lay.ncx;
    P.sz = (long)lay.ncx * lay.ncy;
    P.tiles_x = g.tiles_x;
    P.tiles_y = g.tiles_y;
    P.zchunks = g.zchunks;
    P.kz = g.kz;
    P.nblocks = g.nblocks;
    P.dt = lat.dt;
    P.dt_inv = 1.0 / lat.dt;
    P.etabar = hydro.etabar;
    P.etabar_inv = 1.0 / hydro.etabar;
    P.inv_dx = 1.0 / lat.dx;
    P.inv_dy = 1.0 / lat.dy;
    P.inv_dz = 1.0 / lat.dz;
    P.hx = 1 / lat.dx / 2;
    P.hy = 1 / lat.dy / 2;
    P.hz = 1 / lat.dz / 2;
    P.dt_dx = P.dt * P.inv_dx;
    P.dt_dy = P.dt * P.inv_dy;
    P.dt_dz = P.dt * P.inv_dz;
    P.eos = eos;
    P.counters = counters;
    P.emax = nullptr;
    return P;
}

// The buffers of one Heun step (rungeKutta2, FullyDiscreteKurganovTadmorScheme.cpp:761-789):
// stage 1 reads q^n, e/p^n, u^n, u^{n-1} and writes q*, e/p*, u*; stage 2 reads
// q*, e/p*, u*, u^n and q^n (convex combination) and writes q^{n+1}, e/p^{n+1}
// and u^{n+1}.
struct StepBuffers {
    const double *q_n;
    double *q_star, *q_next;
    const double *e_n, *p_n;
    double *e_star, *p_star, *e_next, *p_next;
    const double *u_n, *u_prev;
    double *u_star, *u_next;
};

void set_stage_buffers(StageParams &P, const StepBuffers &B, int stage, double t, double dt) {
    if (stage == 1) {
        P.qin = B.q_n;
        P.qout = B.q_star;
        P.qn = B.q_n;
        P.e_in = B.e_n;
        P.p_in = B.p_n;
        P.e_out = B.e_star;
        P.p_out = B.p_star;
        P.ucur = B.u_n;
        P.uprev = B.u_prev;
        P.uout = B.u_star;
        P.t_src = t;
        P.t_rec = t + dt;
    } else {
        P.qin = B.q_star;
        P.qout = B.q_next;
        P.qn = B.q_n;
        P.e_in = B.e_star;
        P.p_in = B.p_star;
        P.e_out = B.e_next;
        P.p_out = B.p_next;
        P.ucur = B.u_star;
        P.uprev = B.u_n;
        P.uout = B.u_next;
        P.t_src = t + dt;
        P.t_rec = t + dt;
    }
}

StageParams make_params(const vah_ctx *c, int stage, double t) {
    Geometry g{c->tiles_x, c->tiles_y, c->zchunks, c->kz, c->nblocks, c->grid};
    StageParams P = base_params(c->lat, c->hydro, c->lay, g, c->eos, c->counters);
    StepBuffers B;
    B.q_n = c->qset[c->qi];
    B.q_star = c->qset[c->qs];
    B.q_next = c->qset[c->qo];
    B.e_n = c->ebuf[c->ecur];
    B.p_n = c->pbuf[c->ecur];
    B.e_star = c->ebuf[1 - c->ecur];
    B.p_star = c->pbuf[1 - c->ecur];
    B.e_next = c->ebuf[c->ecur];
    B.p_next = c->pbuf[c->ecur];
    B.u_n = c->useq[c->uc];
    B.u_prev = c->useq[c->up];
    B.u_star = c->useq[c->us];
    B.u_next = c->useq[c->up];
    set_stage_buffers(P, B, stage, t, c->lat.dt);
    return P;
}

// what a stage launch needs besides its parameters
struct LaunchCfg {
    int grid;
    hipStream_t stream;
    int precision, flux_mode;
    bool d3;
};

LaunchCfg launch_cfg(const vah_ctx *c) {
    return LaunchCfg{c->grid, c->stream, c->opt.precision, c->opt.flux_mode, c->lat.nz > 1};
}

template <bool D3, bool S2, bool SH, typename R>
void launch(const LaunchCfg &L, const StageParams &P) {
    hipLaunchKernelGGL((stage_kernel<D3, S2, SH, R>), dim3(L.grid), dim3(NTHREADS), 0, L.stream, P);
}

template <bool D3, bool S2>
void launch_variant(const LaunchCfg &L, const StageParams &P) {
    if (L.precision == VAH_FP32_SPLIT)
        launch<D3, S2, true, float>(L, P);
    else if (L.flux_mode == VAH_FLUX_SHARED)
        launch<D3, S2, true, double>(L, P);
    else
        launch<D3, S2, false, double>(L, P);
}

// 2+1D here; 3+1D in vah_stage3d.hip (compiled with other scheduler settings)
void launch_stage(const LaunchCfg &L, int stage, const StageParams &P) {
    if (L.d3) {
        launch_stage_3d(StageLaunch{L.grid, L.stream, L.precision, L.flux_mode}, stage, P);
    } else {
        if (stage == 1) launch_variant<false, false>(L, P);
        else launch_variant<false, true>(L, P);
    }
}

int fill_ghosts(vah_ctx *c) {
    long n = (long)c->lat.nx * c->lat.ny * c->lat.nz;
    int blocks = (int)((n + 255) / 256);
    hipLaunchKernelGGL(ghost_fill_kernel, dim3(blocks), dim3(256), 0, c->stream, c->qset[c->qi], c->ebuf[c->ecur],
                       c->pbuf[c->ecur], c->useq[c->uc], c->fs, c->lat.nx, c->lat.ny, c->lat.nz, (long)c->ncx,
                       (long)c->ncx * c->ncy, c->lat.nz > 1 ? 1 : 0);
    HIP_TRY(hipGetLastError());
    return VAH_OK;
}

// copy interior host arrays into the padded device layout of one slab
int upload_fields(vah_ctx *c, double *dst_slab, int nfields, const double *src, std::vector<double> &stage_buf) {
    c->lay.pad(nfields, src, stage_buf);
    HIP_TRY(hipMemcpyAsync(dst_slab, stage_buf.data(), sizeof(double) * nfields * c->fs, hipMemcpyHostToDevice,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return VAH_OK;
}

int download_fields(vah_ctx *c, const double *src_slab, int nfields, double *dst) {
    std::vector<double> buf((size_t)nfields * c->fs);
    HIP_TRY(hipMemcpyAsync(buf.data(), src_slab, sizeof(double) * nfields * c->fs, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->lay.unpad(nfields, buf.data(), dst);
    return VAH_OK;
}

__global__ void reduce_kernel(const double *q, const double *e, const double *u, long fs, int nx, int ny, int nz,
                              long sy, long sz, Eos eos, double *out) {
    __shared__ double sm[4][256];
    long n = (long)nx * ny * nz;
    double mx_e = 0.0, mx_T = 0.0, sum = 0.0, mx_v = 0.0;
    for (long m = (long)blockIdx.x * blockDim.x + threadIdx.x; m < n; m += (long)gridDim.x * blockDim.x) {
        int ii = (int)(m % nx);
        int jj = (int)((m / nx) % ny);
        int kk = (int)(m / ((long)nx * ny));
        long s = (long)plane_of(kk, nz > 1) * sz + (long)(jj + 2) * sy + (ii + 2);
        double ev = e[s];
        mx_e = fmax(mx_e, ev);
        mx_T = fmax(mx_T, eos_T(eos, ev));
        sum += q[s];
        double ut = u[s], ux = u[fs + s], uy = u[2 * fs + s];
        mx_v = fmax(mx_v, __builtin_sqrt(ux * ux + uy * uy) / ut);
    }
    int t = threadIdx.x;
    sm[0][t] = mx_e;
    sm[1][t] = mx_T;
    sm[2][t] = sum;
    sm[3][t] = mx_v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            sm[0][t] = fmax(sm[0][t], sm[0][t + w]);
            sm[1][t] = fmax(sm[1][t], sm[1][t + w]);
            sm[2][t] += sm[2][t + w];
            sm[3][t] = fmax(sm[3][t], sm[3][t + w]);
        }
        __syncthreads();
    }
    if (t == 0)
        for (int k = 0; k < 4; ++k) out[blockIdx.x * 4 + k] = sm[k][0];
}

// Second-order inverse Reynolds number of checkValidityKernel
// (HydrodynamicValidity.cpp:106-205 for the shear tensor, :272-387 for Rpi2):
// sigma^{mu nu}, its transverse-traceless part sigma_T, J = delta_pipi
// theta_T pi + tau_pipi I4 - lambda_pipi zDzu pi, and
// sqrt|J:J / sigma_T:sigma_T| / 2 / etabar_TC.  tau_pipi = 4 delta_pipi / 3
// (the reference's integer 1/2 is 0); the bulk term of J multiplies Pi = 0.
// pv: P_L, pi^{tau tau} .. pi^{eta eta} of the step's start.
struct RPi2In {
    double t, ut, ux, uy, un;
    double dtut, dtux, dtuy, dtun, dxut, dxux, dxuy, dxun, dyut, dyux, dyuy, dyun, dnut, dnux, dnuy, dnun;
    double e;
};

__device__ double validity_rpi2(const RPi2In &v, const double (&pv)[11]) {
    const double t = v.t, ut = v.ut, ux = v.ux, uy = v.uy, un = v.un;
    const double t2 = t * t, t3 = t * t2, ut2 = ut * ut, un2 = un * un, it = 1.0 / t, it2 = it * it;
    const double dut = ut * v.dtut + ux * v.dxut + uy * v.dyut + un * v.dnut;
    const double dux = ut * v.dtux + ux * v.dxux + uy * v.dyux + un * v.dnux;
    const double duy = ut * v.dtuy + ux * v.dxuy + uy * v.dyuy + un * v.dnuy;
    const double dun = ut * v.dtun + ux * v.dxun + uy * v.dyun + un * v.dnun;
    const double theta = ut * it + v.dtut + v.dxux + v.dyuy + v.dnun;
    const double th3 = theta / 3;
    // sigma^{mu nu}
    const double stt = -t * ut * un2 + (v.dtut - ut * dut) + (ut2 - 1) * th3;
    const double stx = 0.5 * (-t * un2 * ux + (v.dtux - v.dxut) - (ux * dut + ut * dux)) + ut * ux * th3;
    const double sty = 0.5 * (-t * un2 * uy + (v.dtuy - v.dyut) - (uy * dut + ut * duy)) + ut * uy * th3;
    const double stn = -un * (2 * ut2 + t2 * un2) * 0.5 * it + 0.5 * (v.dtun - v.dnut * it2) -
                       0.5 * (un * dut + ut * dun) + ut * un * th3;
    const double sxx = -(v.dxux + ux * dux) + (1 + ux * ux) * th3;
    const double sxy = -0.5 * (v.dxuy + v.dyux) - 0.5 * (uy * dux + ux * duy) + ux * uy * th3;
    const double sxn = -ut * ux * un * it - 0.5 * (v.dxun + v.dnux * it2) - 0.5 * (un * dux + ux * dun) + ux * un * th3;
    const double syy = -(v.dyuy + uy * duy) + (1 + uy * uy) * th3;
    const double syn = -ut * uy * un * it - 0.5 * (v.dyun + v.dnuy * it2) - 0.5 * (un * duy + uy * dun) + uy * un * th3;
    const double snn = -ut * (1 + 2 * t2 * un2) / t3 - v.dnun * it2 - un * dun + (it2 + un2) * th3;
    // z, A, B and the transverse projections
    const double iF = vah::rsqrt(1 + ux * ux + uy * uy);
    const double z0 = t * un * iF, z3 = ut * it * iF;
    const double z02 = z0 * z0, z0z3 = z0 * z3, z32 = z3 * z3;
    const double A = z02 * stt - 2 * t2 * z0z3 * stn + t2 * t2 * z32 * snn;
    const double B0 = z0 * stt - t2 * z3 * stn, B1 = z0 * stx - t2 * z3 * sxn;
    const double B2 = z0 * sty - t2 * z3 * syn, B3 = z0 * stn - t2 * z3 * snn;
    const double thetaT = 2 * th3 + A;
    const double zDzu = theta - thetaT;
    const double hA = 0.5 * A;
    const double sttT = stt + 2 * z0 * B0 + z02 * A - (1 - ut2 + z02) * hA;
    const double stxT = stx + z0 * B1 + ut * ux * hA;
    const double styT = sty + z0 * B2 + ut * uy * hA;
    const double stnT = stn + z0 * B3 + z3 * B0 + z0z3 * A + (ut * un - z0z3) * hA;
    const double sxxT = sxx + (1 + ux * ux) * hA;
    const double sxyT = sxy + ux * uy * hA;
    const double sxnT = sxn + z3 * B1 + ux * un * hA;
    const double syyT = syy + (1 + uy * uy) * hA;
    const double synT = syn + z3 * B2 + uy * un * hA;
    const double snnT = snn + 2 * z3 * B3 + z32 * A + (it2 + un2 - z32) * hA;
    const double e_s = v.e, pl = pv[0];
    const double pitt = pv[1], pitx = pv[2], pity = pv[3], pitn = pv[4], pixx = pv[5], pixy = pv[6], pixn = pv[7];
    const double piyy = pv[8], piyn = pv[9], pinn = pv[10];
    const Fits f = fits(powers_of(pl / e_s));
    const double etaBar_TC = ((3 - f.Rtilde) * e_s - 2 * pl) * 0.125;
    const double delta_pipi = (3 + f.Rgamma) * 0.5;
    const double tau_pipi = 4 * delta_pipi / 3;
    const double lambda_pipi = f.Rgamma - 1;
    const double ps2 = 0.5 * (pitt * sttT - 2 * pitx * stxT - 2 * pity * styT + pixx * sxxT + 2 * pixy * sxyT +
                              piyy * syyT + 2 * t2 * (pixn * sxnT + piyn * synT - pitn * stnT) + pinn * snnT * t2 * t2);
    const double I4tt = (pitt * sttT - pitx * stxT - pity * styT - t2 * pitn * stnT) - (1 - ut2 + z02) * ps2;
    const double I4tx = 0.5 * ((pitt * stxT + pitx * sttT) - (pitx * sxxT + pixx * stxT) - (pity * sxyT + pixy * styT) -
                               t2 * (pitn * sxnT + pixn * stnT)) + ut * ux * ps2;
    const double I4ty = 0.5 * ((pitt * styT + pity * sttT) - (pitx * sxyT + pixy * stxT) - (pity * syyT + piyy * styT) -
                               t2 * (pitn * synT + piyn * stnT)) + ut * uy * ps2;
    const double I4tn = 0.5 * ((pitt * stnT + pitn * sttT) - (pitx * sxnT + pixn * stxT) - (pity * synT + piyn * styT) -
                               t2 * (pitn * snnT + pinn * stnT)) + (ut * un - z0z3) * ps2;
    const double I4xx = (pitx * stxT - pixx * sxxT - pixy * sxyT - t2 * pixn * sxnT) + (1 + ux * ux) * ps2;
    const double I4xy = 0.5 * ((pitx * styT + pity * stxT) - (pixx * sxyT + pixy * sxxT) - (pixy * syyT + piyy * sxyT) -
                               t2 * (pixn * synT + piyn * sxnT)) + ux * uy * ps2;
    const double I4xn = 0.5 * ((pitx * stnT + pitn * stxT) - (pixx * sxnT + pixn * sxxT) - (pixy * synT + piyn * sxyT) -
                               t2 * (pixn * snnT + pinn * sxnT)) + ux * un * ps2;
    const double I4yy = (pity * styT - pixy * sxyT - piyy * syyT - t2 * piyn * synT) + (1 + uy * uy) * ps2;
    const double I4yn = 0.5 * ((pity * stnT + pitn * styT) - (pixy * sxnT + pixn * sxyT) - (piyy * synT + piyn * syyT) -
                               t2 * (piyn * snnT + pinn * synT)) + uy * un * ps2;
    const double I4nn = (pitn * stnT - pixn * sxnT - piyn * synT - t2 * pinn * snnT) + (it2 + un2 - z32) * ps2;
    const double c2 = delta_pipi * thetaT - lambda_pipi * zDzu;
    const double Jtt = c2 * pitt + tau_pipi * I4tt, Jtx = c2 * pitx + tau_pipi * I4tx;
    const double Jty = c2 * pity + tau_pipi * I4ty, Jtn = c2 * pitn + tau_pipi * I4tn;
    const double Jxx = c2 * pixx + tau_pipi * I4xx, Jxy = c2 * pixy + tau_pipi * I4xy;
    const double Jxn = c2 * pixn + tau_pipi * I4xn, Jyy = c2 * piyy + tau_pipi * I4yy;
    const double Jyn = c2 * piyn + tau_pipi * I4yn, Jnn = c2 * pinn + tau_pipi * I4nn;
    auto contract = [t2](double tt, double tx, double ty, double tn, double xx, double xy, double xn, double yy,
                         double yn, double nn) {
        return tt * tt + xx * xx + yy * yy + 2 * (xy * xy - tx * tx - ty * ty) +
               2 * t2 * (xn * xn + yn * yn - tn * tn) + nn * nn * t2 * t2;
    };
    const double ss = contract(sttT, stxT, styT, stnT, sxxT, sxyT, sxnT, syyT, synT, snnT);
    const double JJ = contract(Jtt, Jtx, Jty, Jtn, Jxx, Jxy, Jxn, Jyy, Jyn, Jnn);
    return __builtin_sqrt(fabs(JJ / ss)) * 0.5 / etaBar_TC;
}

// Validity diagnostics (HydrodynamicValidity.cpp:47-387) with the reference's
// arguments after a step: t_{n+1} = t_n + dt (rungeKutta2 advances t before
// its checkValidity call), the conserved variables q^n of the step's start,
// the new e, p, u and u^n as the previous velocity.
// out[f*n + m] for f = kn_T, kn_L, kn, r_pi, r_pi2 (m: interior cell, x fastest).
__global__ void validity_kernel(const double *qn, const double *e, const double *p, const double *u,
                                const double *up, long fs, int nx, int ny, int nz, long sy, long sz, double t,
                                double dt, double etabar, double facX, double facY, double facZ, Eos eos,
                                double *out) {
    const long n = (long)nx * ny * nz;
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const int ii = (int)(m % nx);
    const int jj = (int)((m / nx) % ny);
    const int kk = (int)(m / ((long)nx * ny));
    const long s = (long)plane_of(kk, nz > 1) * sz + (long)(jj + 2) * sy + (ii + 2);
    RPi2In v;
    v.t = t;
    v.e = e[s];
    double uc[4], d[4][4];  // d[axis][component]: axis 0 = tau, 1 = x, 2 = y, 3 = eta
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const double *f = u + c * fs;
        uc[c] = f[s];
        d[0][c] = (uc[c] - up[c * fs + s]) / dt;
        d[1][c] = (f[s + 1] - f[s - 1]) * facX;
        d[2][c] = (f[s + sy] - f[s - sy]) * facY;
        d[3][c] = nz > 1 ? (f[s + sz] - f[s - sz]) * facZ : 0.0;
    }
    v.ut = uc[0];
    v.ux = uc[1];
    v.uy = uc[2];
    v.un = uc[3];
    v.dtut = d[0][0]; v.dtux = d[0][1]; v.dtuy = d[0][2]; v.dtun = d[0][3];
    v.dxut = d[1][0]; v.dxux = d[1][1]; v.dxuy = d[1][2]; v.dxun = d[1][3];
    v.dyut = d[2][0]; v.dyux = d[2][1]; v.dyuy = d[2][2]; v.dyun = d[2][3];
    v.dnut = d[3][0]; v.dnux = d[3][1]; v.dnuy = d[3][2]; v.dnun = d[3][3];
    const double ut = v.ut, ux = v.ux, uy = v.uy, un = v.un;
    const double ut2 = ut * ut, un2 = un * un, t2 = t * t, t3 = t * t2;
    const double Dut = ut * v.dtut + ux * v.dxut + uy * v.dyut + un * v.dnut + t * un * un;
    const double dut = Dut - t * un * un;
    const double dun = ut * v.dtun + ux * v.dxun + uy * v.dyun + un * v.dnun;
    const double theta = ut / t + v.dtut + v.dxux + v.dyuy + v.dnun;
    const double stt = -t * ut * un2 + (v.dtut - ut * dut) + (ut2 - 1) * theta / 3;
    const double stn = -un * (2 * ut2 + t2 * un2) / (2 * t) + (v.dtun - v.dnut / t2) / 2 - (un * dut + ut * dun) / 2 +
                       ut * un * theta / 3;
    const double snn = -ut * (1 + 2 * t2 * un2) / t3 - v.dnun / t2 - un * dun + (1 / t2 + un2) * theta / 3;
    const double F = 1 + ux * ux + uy * uy;
    const double z0 = t * un / __builtin_sqrt(F), z3 = ut / t / __builtin_sqrt(F);
    const double A = z0 * z0 * stt - 2 * t2 * z0 * z3 * stn + t2 * t2 * z3 * z3 * snn;
    const double thetaT = 2 * theta / 3 + A;
    const double zDzu = theta / 3 - A;
    const double T = eos_T(eos, v.e);
    out[m] = 5 * etabar * fabs(thetaT) / T;
    out[n + m] = 5 * etabar * fabs(zDzu) / T;
    out[2 * n + m] = 5 * etabar * fabs(theta) / T;
    double pv[11];
#pragma unroll
    for (int f = 0; f < 11; ++f) pv[f] = qn[(PL + f) * fs + s];
    const double pitt = pv[1], pitx = pv[2], pity = pv[3], pitn = pv[4], pixx = pv[5], pixy = pv[6];
    const double pixn = pv[7], piyy = pv[8], piyn = pv[9], pinn = pv[10];
    const double pipi = pitt * pitt - 2 * pitx * pitx - 2 * pity * pity + pixx * pixx + 2 * pixy * pixy + piyy * piyy -
                        2 * pitn * pitn * t2 + 2 * pixn * pixn * t2 + 2 * piyn * piyn * t2 + pinn * pinn * t2 * t2;
    out[3 * n + m] = __builtin_sqrt(fabs(pipi)) / fabs(pt_hat(v.e, p[s], pv[0]));
    out[4 * n + m] = validity_rpi2(v, pv);
}

// ---- Scratch of the no-allocation step ------------------------------------
// One file-scope device array per GPU holds the intermediate stage of
// vah_step_noalloc (q*, e*, p*, u* of the padded lattice) and its NaN /
// fallback counters; a second one the conformal EOS's powf bitmap.  Both
// exist on every device the library is loaded on until the process ends.
// Host code never uses their names as addresses: the device addresses come
// from hipGetSymbolAddress, once per device, with that device current.
constexpr size_t kScratchDoubles = VAH_NOALLOC_SCRATCH_BYTES / sizeof(double);
constexpr int kStageFields = NQ + 2 + 4;  // q*, e*, p*, u*
static __device__ __align__(256) double scratch[kScratchDoubles];
static __device__ __align__(256) uint32_t powf_fix_static[1u << 20];  // 2^25 bits

constexpr int kMaxDevices = 64;
struct NoallocDevice {
    std::once_flag once;
    int rc = VAH_OK;
    std::string err;
    double *scratch = nullptr;      // device address of `scratch` on this device
    uint32_t *powf_fix = nullptr;   // device address of `powf_fix_static`
};
NoallocDevice g_noalloc[kMaxDevices];

// First call on a device (that device current): the two symbol addresses and
// the powf bitmap upload.  Synchronous; later calls only enqueue work.
int noalloc_setup(int dev, NoallocDevice **out) {
    NoallocDevice &d = g_noalloc[dev];
    std::call_once(d.once, [&d] {
        hipError_t e = hipGetSymbolAddress((void **)&d.scratch, HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            d.rc = VAH_ERR_HIP;
            d.err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return;
        }
        e = hipGetSymbolAddress((void **)&d.powf_fix, HIP_SYMBOL(powf_fix_static));
        if (e != hipSuccess) {
            d.rc = VAH_ERR_HIP;
            d.err = std::string("hipGetSymbolAddress(powf_fix_static): ") + hipGetErrorString(e);
            return;
        }
        const std::vector<uint32_t> &fix = vah_host::powf_fix_table();
        e = hipMemcpy(d.powf_fix, fix.data(), fix.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            d.rc = VAH_ERR_HIP;
            d.err = std::string("powf bitmap upload: ") + hipGetErrorString(e);
        }
    });
    if (d.rc != VAH_OK) return fail(d.rc, "vah_step_noalloc: set-up on device %d: %s", dev, d.err.c_str());
    *out = &d;
    return VAH_OK;
}

__global__ void clear_counters_kernel(unsigned int *c) {
    if (threadIdx.x < 4) c[threadIdx.x] = 0u;
}

}  // namespace

extern "C" {

void vah_default_options(vah_options *opt) {
    if (!opt) return;
    memset(opt, 0, sizeof *opt);
    opt->eos = VAH_EOS_CONFORMAL;
    opt->precision = VAH_FP64;
    opt->flux_mode = VAH_FLUX_SHARED;
    opt->device = 0;
    opt->stream = nullptr;
}

const char *vah_last_error(void) { return g_err.c_str(); }

const char *vah_build_info(void) { return "vah-mi355x gfx950 fused-KT-stage fp64 fp32-split"; }

int vah_create(const vah_lattice *lat, const vah_hydro *hydro, const vah_options *opt_in, vah_ctx **out) {
    if (!out) return fail(VAH_ERR_ARG, "vah_create: out is NULL");
    *out = nullptr;
    if (!lattice_ok(lat)) return fail(VAH_ERR_ARG, "vah_create: lattice sizes and spacings must be positive");
    if (!hydro || !(hydro->t0 > 0) || !(hydro->etabar > 0))
        return fail(VAH_ERR_ARG, "vah_create: need t0 > 0 and eta/s > 0");
    vah_options opt;
    if (opt_in)
        opt = *opt_in;
    else
        vah_default_options(&opt);
    if (opt.precision != VAH_FP64 && opt.precision != VAH_FP32_SPLIT)
        return fail(VAH_ERR_ARG, "vah_create: unknown precision %d", opt.precision);
    if (opt.precision == VAH_FP32_SPLIT && opt.flux_mode != VAH_FLUX_SHARED)
        return fail(VAH_ERR_ARG, "vah_create: the split-precision form uses shared faces only");
    if (opt.eta_chunk < 0 || opt.eta_chunk > 256)
        return fail(VAH_ERR_ARG, "vah_create: eta_chunk must be in [0, 256]");
    if (opt.flux_mode != VAH_FLUX_SHARED && opt.flux_mode != VAH_FLUX_PER_CELL)
        return fail(VAH_ERR_ARG, "vah_create: unknown flux mode %d", opt.flux_mode);
    if (opt.eos == VAH_EOS_TABLE && (opt.table.n < 2 || !opt.table.p || !opt.table.T || !(opt.table.h > 0)))
        return fail(VAH_ERR_ARG, "vah_create: EOS table needs n >= 2 nodes, h > 0 and p, T arrays");
    if (opt.eos != VAH_EOS_CONFORMAL && opt.eos != VAH_EOS_TABLE && opt.eos != VAH_EOS_WB)
        return fail(VAH_ERR_ARG, "vah_create: unknown EOS kind %d", opt.eos);

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= opt.device || opt.device < 0)
        return fail(VAH_ERR_DEVICE, "vah_create: HIP device %d not available (%d found)", opt.device, ndev);
    HIP_TRY(hipSetDevice(opt.device));

    {
        // each 15-field set is addressed through one buffer resource (< 4 GB)
        vah_host::Layout L(lat->nx, lat->ny, lat->nz);
        if ((double)L.fs * NQ * 8 >= 4294967296.0)
            return fail(VAH_ERR_ARG, "vah_create: %ld padded cells per field exceed one context's 4 GB field-set "
                        "limit; split the lattice in eta slabs", L.fs);
    }
    vah_ctx *c = new vah_ctx();
    c->lat = *lat;
    c->hydro = *hydro;
    c->opt = opt;
    c->device = opt.device;
    c->lay = vah_host::Layout(lat->nx, lat->ny, lat->nz);
    c->ncx = c->lay.ncx;
    c->ncy = c->lay.ncy;
    c->ncz = c->lay.ncz;
    c->ncell = c->lay.ncell;
    c->fs = c->lay.fs;
    c->qi = 0;
    c->qs = 1;
    c->qo = 2;
    c->uc = 0;
    c->up = 1;
    c->us = 2;
    c->ecur = 0;

    auto cleanup = [&](int code) {
        vah_destroy(c);
        return code;
    };
    if (opt.stream) {
        c->stream = (hipStream_t)opt.stream;
        c->own_stream = false;
    } else {
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
            return cleanup(fail(VAH_ERR_HIP, "vah_create: stream creation failed"));
        c->own_stream = true;
    }
    const size_t nslab = 3 * NQ + 4 + 12;
    size_t bytes = sizeof(double) * c->fs * nslab;
    if (hipMalloc(&c->slab, bytes) != hipSuccess)
        return cleanup(fail(VAH_ERR_NOMEM, "vah_create: cannot allocate %.2f GB of device state", bytes / 1e9));
    if (hipMemsetAsync(c->slab, 0, bytes, c->stream) != hipSuccess)
        return cleanup(fail(VAH_ERR_HIP, "vah_create: memset failed"));
    double *b = c->slab;
    for (int k = 0; k < 3; ++k, b += NQ * c->fs) c->qset[k] = b;
    for (int k = 0; k < 2; ++k, b += c->fs) c->ebuf[k] = b;
    for (int k = 0; k < 2; ++k, b += c->fs) c->pbuf[k] = b;
    for (int k = 0; k < 3; ++k, b += 4 * c->fs) c->useq[k] = b;

    if (hipMalloc(&c->counters, 4 * sizeof(unsigned)) != hipSuccess ||
        hipMemsetAsync(c->counters, 0, 4 * sizeof(unsigned), c->stream) != hipSuccess)
        return cleanup(fail(VAH_ERR_NOMEM, "vah_create: counters"));

    {
        std::string err;
        int rc = vah_host::make_device_eos(opt.eos, opt.table, c->stream, &c->eos, &c->powf_fix, &c->eos_dev,
                                           &c->heos, &err);
        if (rc != VAH_OK) return cleanup(fail(rc, "vah_create: %s", err.c_str()));
        // the table arrays belong to the caller: only the copies are kept
        c->opt.table.p = nullptr;
        c->opt.table.T = nullptr;
    }

    // launch geometry
    {
        Geometry g = make_geometry(*lat, opt.eta_chunk);
        c->tiles_x = g.tiles_x;
        c->tiles_y = g.tiles_y;
        c->kz = g.kz;
        c->zchunks = g.zchunks;
        c->nblocks = g.nblocks;
        c->grid = g.grid;
    }
    c->red_blocks = 1024;
    if (hipMalloc(&c->red_buf, sizeof(double) * 4 * c->red_blocks) != hipSuccess)
        return cleanup(fail(VAH_ERR_NOMEM, "vah_create: reduction buffer"));
    if (hipStreamSynchronize(c->stream) != hipSuccess) return cleanup(fail(VAH_ERR_HIP, "vah_create: sync"));
    *out = c;
    return VAH_OK;
}

void vah_destroy(vah_ctx *c) {
    if (!c) return;
    // teardown: errors are not actionable here
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &e : c->ev) (void)hipEventDestroy(e);
    (void)hipFree(c->slab);
    (void)hipFree(c->counters);
    (void)hipFree(c->powf_fix);
    (void)hipFree(c->eos_dev);
    (void)hipFree(c->red_buf);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int vah_set_primitives(vah_ctx *c, const double *e, const double *pl, const double *ut, const double *ux,
                       const double *uy, const double *un, const double *pi) {
    if (!c || !e || !ut || !ux || !uy || !un) return fail(VAH_ERR_ARG, "vah_set_primitives: NULL argument");
    HIP_TRY(hipSetDevice(c->device));
    const long n = (long)c->lat.nx * c->lat.ny * c->lat.nz;
    std::vector<double> q((size_t)NQ * n), ev(n), pv(n), u((size_t)4 * n);
    vah_host::prims_to_conserved(c->hydro.t0, c->heos, n, e, pl, ut, ux, uy, un, pi, q.data(), ev.data(), pv.data(),
                                 u.data());
    return vah_set_state(c, 0, q.data(), ev.data(), pv.data(), u.data(), u.data());
}

int vah_set_state(vah_ctx *c, long long nstep, const double *q, const double *e, const double *p, const double *u,
                  const double *up) {
    if (!c || !q || !e || !p || !u || !up) return fail(VAH_ERR_ARG, "vah_set_state: NULL argument");
    if (nstep < 0) return fail(VAH_ERR_ARG, "vah_set_state: negative step count");
    HIP_TRY(hipSetDevice(c->device));
    std::vector<double> buf;
    int rc;
    if ((rc = upload_fields(c, c->qset[c->qi], NQ, q, buf)) != VAH_OK) return rc;
    if ((rc = upload_fields(c, c->ebuf[c->ecur], 1, e, buf)) != VAH_OK) return rc;
    if ((rc = upload_fields(c, c->pbuf[c->ecur], 1, p, buf)) != VAH_OK) return rc;
    if ((rc = upload_fields(c, c->useq[c->uc], 4, u, buf)) != VAH_OK) return rc;
    if ((rc = upload_fields(c, c->useq[c->up], 4, up, buf)) != VAH_OK) return rc;
    if ((rc = fill_ghosts(c)) != VAH_OK) return rc;
    // NaN and fallback counts restart with the new state
    HIP_TRY(hipMemsetAsync(c->counters, 0, 4 * sizeof(unsigned), c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->counters_seen[0] = c->counters_seen[1] = 0;
    c->nstep = nstep;
    c->have_state = true;
    c->stepped = false;
    return VAH_OK;
}

int vah_step(vah_ctx *c, int nsteps) {
    if (!c) return fail(VAH_ERR_ARG, "vah_step: NULL context");
    if (!c->have_state) return fail(VAH_ERR_STATE, "vah_step: no initial state");
    if (nsteps < 0) return fail(VAH_ERR_ARG, "vah_step: negative step count");
    HIP_TRY(hipSetDevice(c->device));
    for (int it = 0; it < nsteps; ++it) {
        // tau_n = t0 + n dt, as the driver recomputes it (HydroPlugin.cpp:232)
        const double t = c->hydro.t0 + (double)c->nstep * c->lat.dt;
        for (int stage = 1; stage <= 2; ++stage) {
            StageParams P = make_params(c, stage, t);
            if (c->prof) {
                hipEvent_t a, b;
                HIP_TRY(hipEventCreate(&a));
                HIP_TRY(hipEventCreate(&b));
                HIP_TRY(hipEventRecord(a, c->stream));
                launch_stage(launch_cfg(c), stage, P);
                HIP_TRY(hipEventRecord(b, c->stream));
                c->ev.push_back(a);
                c->ev.push_back(b);
                c->ev_kind.push_back(stage - 1);
            } else {
                launch_stage(launch_cfg(c), stage, P);
            }
            HIP_TRY(hipGetLastError());
        }
        // rotate roles: q^{n+1} becomes current; (u^n, u^{n-1}) <- (new, old u^n)
        std::swap(c->qi, c->qo);
        int old_uc = c->uc;
        c->uc = c->up;
        c->up = old_uc;
        c->nstep++;
        c->stepped = true;
    }
    return VAH_OK;
}

int vah_synchronize(vah_ctx *c) {
    if (!c) return fail(VAH_ERR_ARG, "vah_synchronize: NULL context");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    unsigned h[4];
    HIP_TRY(hipMemcpy(h, c->counters, sizeof h, hipMemcpyDeviceToHost));
    if (h[0] != c->counters_seen[0]) {
        unsigned d = h[0] - c->counters_seen[0];
        c->counters_seen[0] = h[0];
        c->counters_seen[1] = h[1];
        return fail(VAH_ERR_NAN, "vah_synchronize: %u cell updates produced NaN", d);
    }
    c->counters_seen[1] = h[1];
    return VAH_OK;
}

int vah_get_state(vah_ctx *c, double *q, double *e, double *p, double *u, double *up) {
    if (!c) return fail(VAH_ERR_ARG, "vah_get_state: NULL context");
    HIP_TRY(hipSetDevice(c->device));
    int rc;
    if (q && (rc = download_fields(c, c->qset[c->qi], NQ, q)) != VAH_OK) return rc;
    if (e && (rc = download_fields(c, c->ebuf[c->ecur], 1, e)) != VAH_OK) return rc;
    if (p && (rc = download_fields(c, c->pbuf[c->ecur], 1, p)) != VAH_OK) return rc;
    if (u && (rc = download_fields(c, c->useq[c->uc], 4, u)) != VAH_OK) return rc;
    if (up && (rc = download_fields(c, c->useq[c->up], 4, up)) != VAH_OK) return rc;
    return VAH_OK;
}

int vah_get_time(const vah_ctx *c, double *tau, long long *nstep) {
    if (!c) return fail(VAH_ERR_ARG, "vah_get_time: NULL context");
    if (tau) *tau = c->hydro.t0 + (double)c->nstep * c->lat.dt;
    if (nstep) *nstep = c->nstep;
    return VAH_OK;
}

int vah_reduce(vah_ctx *c, vah_stats *out) {
    if (!c || !out) return fail(VAH_ERR_ARG, "vah_reduce: NULL argument");
    HIP_TRY(hipSetDevice(c->device));
    hipLaunchKernelGGL(reduce_kernel, dim3(c->red_blocks), dim3(256), 0, c->stream, c->qset[c->qi], c->ebuf[c->ecur],
                       c->useq[c->uc], c->fs, c->lat.nx, c->lat.ny, c->lat.nz, (long)c->ncx, (long)c->ncx * c->ncy,
                       c->eos, c->red_buf);
    HIP_TRY(hipGetLastError());
    std::vector<double> h((size_t)4 * c->red_blocks);
    HIP_TRY(hipMemcpyAsync(h.data(), c->red_buf, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    double mx_e = 0, mx_T = 0, sum = 0, mx_v = 0;
    for (int b = 0; b < c->red_blocks; ++b) {
        mx_e = std::max(mx_e, h[4 * b]);
        mx_T = std::max(mx_T, h[4 * b + 1]);
        sum += h[4 * b + 2];
        mx_v = std::max(mx_v, h[4 * b + 3]);
    }
    unsigned cnt[4];
    HIP_TRY(hipMemcpy(cnt, c->counters, sizeof cnt, hipMemcpyDeviceToHost));
    const double tau = c->hydro.t0 + (double)c->nstep * c->lat.dt;
    out->max_e = mx_e;
    out->max_T = mx_T;
    out->total_energy = tau * sum * c->lat.dx * c->lat.dy * c->lat.dz;
    out->max_vT = mx_v;
    out->recovery_fallbacks = cnt[1];
    out->nan_cells = cnt[0];
    return VAH_OK;
}

int vah_validity(vah_ctx *c, double *kn_T, double *kn_L, double *kn, double *r_pi, double *r_pi2) {
    if (!c) return fail(VAH_ERR_ARG, "vah_validity: NULL context");
    if (!c->stepped)
        return fail(VAH_ERR_STATE, "vah_validity: needs a step completed since the last state upload");
    HIP_TRY(hipSetDevice(c->device));
    const long n = (long)c->lat.nx * c->lat.ny * c->lat.nz;
    // the stage-1 output set is free between steps: 5 interior-sized outputs fit in it
    double *out = c->qset[c->qs];
    // the reference's t after its step: t_n, as the driver computes it, plus dt
    // (FullyDiscreteKurganovTadmorScheme.cpp:766,791)
    const double t = (c->hydro.t0 + (double)(c->nstep - 1) * c->lat.dt) + c->lat.dt;
    hipLaunchKernelGGL(validity_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->qset[c->qo],
                       c->ebuf[c->ecur], c->pbuf[c->ecur], c->useq[c->uc], c->useq[c->up], c->fs, c->lat.nx,
                       c->lat.ny, c->lat.nz, (long)c->ncx, (long)c->ncx * c->ncy, t, c->lat.dt, c->hydro.etabar,
                       1 / c->lat.dx / 2, 1 / c->lat.dy / 2, 1 / c->lat.dz / 2, c->eos, out);
    HIP_TRY(hipGetLastError());
    double *dst[5] = {kn_T, kn_L, kn, r_pi, r_pi2};
    for (int f = 0; f < 5; ++f)
        if (dst[f])
            HIP_TRY(hipMemcpyAsync(dst[f], out + (size_t)f * n, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return VAH_OK;
}

int vah_step_noalloc(const vah_lattice *lat, const vah_hydro *hydro, const vah_options *opt, long long nstep,
                     int nsteps, vah_device_state *st, unsigned int *counters) {
    if (!lattice_ok(lat)) return fail(VAH_ERR_ARG, "vah_step_noalloc: lattice sizes and spacings must be positive");
    if (!hydro || !(hydro->t0 > 0) || !(hydro->etabar > 0))
        return fail(VAH_ERR_ARG, "vah_step_noalloc: need t0 > 0 and eta/s > 0");
    if (!opt || !st) return fail(VAH_ERR_ARG, "vah_step_noalloc: NULL options or state");
    if (nstep < 0 || nsteps < 0) return fail(VAH_ERR_ARG, "vah_step_noalloc: negative step count");
    if (!st->q || !st->q_next || !st->e || !st->p || !st->u || !st->u_prev)
        return fail(VAH_ERR_ARG, "vah_step_noalloc: NULL state buffer");
    if (opt->precision != VAH_FP64 && opt->precision != VAH_FP32_SPLIT)
        return fail(VAH_ERR_ARG, "vah_step_noalloc: unknown precision %d", opt->precision);
    if (opt->flux_mode != VAH_FLUX_SHARED && opt->flux_mode != VAH_FLUX_PER_CELL)
        return fail(VAH_ERR_ARG, "vah_step_noalloc: unknown flux mode %d", opt->flux_mode);
    if (opt->precision == VAH_FP32_SPLIT && opt->flux_mode != VAH_FLUX_SHARED)
        return fail(VAH_ERR_ARG, "vah_step_noalloc: the split-precision form uses shared faces only");
    if (opt->eta_chunk < 0 || opt->eta_chunk > 256)
        return fail(VAH_ERR_ARG, "vah_step_noalloc: eta_chunk must be in [0, 256]");
    if (opt->eos == VAH_EOS_TABLE &&
        (opt->table.n < 2 || !opt->table.p || !opt->table.T || !(opt->table.h > 0)))
        return fail(VAH_ERR_ARG, "vah_step_noalloc: EOS table needs n >= 2 nodes, h > 0 and device p, T arrays");
    if (opt->eos != VAH_EOS_CONFORMAL && opt->eos != VAH_EOS_TABLE && opt->eos != VAH_EOS_WB)
        return fail(VAH_ERR_ARG, "vah_step_noalloc: unknown EOS kind %d", opt->eos);
    const vah_host::Layout lay(lat->nx, lat->ny, lat->nz);
    const size_t need = (size_t)kStageFields * (size_t)lay.fs + 64;  // + 64 doubles of counters
    if (need > kScratchDoubles)
        return fail(VAH_ERR_ARG, "vah_step_noalloc: %ld padded cells per field exceed the scratch limit of %ld "
                    "(%zu MiB for %d fields); use a context (vah_create)", lay.fs,
                    (long)((kScratchDoubles - 64) / kStageFields), (size_t)(VAH_NOALLOC_SCRATCH_BYTES >> 20),
                    kStageFields);
    int dev = 0;
    HIP_TRY(hipSetDevice(opt->device));
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(VAH_ERR_DEVICE, "vah_step_noalloc: device %d out of range", dev);
    NoallocDevice *D = nullptr;
    int rc = noalloc_setup(dev, &D);
    if (rc != VAH_OK) return rc;
    if (nsteps == 0) return VAH_OK;

    // the EOS as the stage kernels read it: table arrays are the caller's device
    // memory, the conformal bitmap is the static copy
    Eos eos;
    memset(&eos, 0, sizeof eos);
    eos.kind = opt->eos;
    if (opt->eos == VAH_EOS_CONFORMAL) {
        eos.powf_fix = D->powf_fix;
    } else if (opt->eos == VAH_EOS_TABLE) {
        eos.n = opt->table.n;
        eos.lne0 = opt->table.lne0;
        eos.inv_h = 1.0 / opt->table.h;
        eos.e_first = exp(opt->table.lne0);
        eos.e_last = exp(opt->table.lne0 + (opt->table.n - 1) / (1.0 / opt->table.h));
        eos.p = opt->table.p;
        eos.T = opt->table.T;
    }
    double *scr = D->scratch;
    unsigned int *cnt = counters ? counters : (unsigned int *)(scr + (size_t)kStageFields * lay.fs);
    hipStream_t stream = (hipStream_t)opt->stream;
    if (!counters) {
        // a fill kernel (no memset node when the caller captures the stream)
        hipLaunchKernelGGL(clear_counters_kernel, dim3(1), dim3(64), 0, stream, cnt);
        HIP_TRY(hipGetLastError());
    }
    // zero-gradient ghosts of the inputs (every later input is a stage output,
    // whose ghosts the stage kernels write)
    {
        const long n = lay.interior();
        hipLaunchKernelGGL(ghost_fill_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, st->q, st->e,
                           st->p, st->u, lay.fs, lat->nx, lat->ny, lat->nz, (long)lay.ncx,
                           (long)lay.ncx * lay.ncy, lat->nz > 1 ? 1 : 0);
        HIP_TRY(hipGetLastError());
    }
    const Geometry g = make_geometry(*lat, opt->eta_chunk);
    const LaunchCfg L{g.grid, stream, opt->precision, opt->flux_mode, lat->nz > 1};
    StageParams P = base_params(*lat, *hydro, lay, g, eos, cnt);
    StepBuffers B;
    B.q_star = scr;
    B.e_star = scr + (size_t)NQ * lay.fs;
    B.p_star = scr + (size_t)(NQ + 1) * lay.fs;
    B.u_star = scr + (size_t)(NQ + 2) * lay.fs;
    for (int it = 0; it < nsteps; ++it, ++nstep) {
        // tau_n = t0 + n dt, as vah_step and the reference's driver compute it
        const double t = hydro->t0 + (double)nstep * lat->dt;
        B.q_n = st->q;
        B.q_next = st->q_next;
        B.e_n = st->e;
        B.p_n = st->p;
        B.e_next = st->e;  // stage 2 reads e*, p* only: e, p are updated in place
        B.p_next = st->p;
        B.u_n = st->u;
        B.u_prev = st->u_prev;
        B.u_next = st->u_prev;  // u^{n-1} is dead after stage 1
        for (int stage = 1; stage <= 2; ++stage) {
            set_stage_buffers(P, B, stage, t, lat->dt);
            launch_stage(L, stage, P);
            HIP_TRY(hipGetLastError());
        }
        // the reference's swaps: q <-> Q (HydroPlugin.cpp:230), up <-> u
        std::swap(st->q, st->q_next);
        std::swap(st->u, st->u_prev);
    }
    return VAH_OK;
}

int vah_profile_enable(vah_ctx *c, int on) {
    if (!c) return fail(VAH_ERR_ARG, "vah_profile_enable: NULL context");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (auto &e : c->ev) (void)hipEventDestroy(e);
    c->ev.clear();
    c->ev_kind.clear();
    c->prof = on != 0;
    return VAH_OK;
}

int vah_profile_read(vah_ctx *c, vah_profile *out) {
    if (!c || !out) return fail(VAH_ERR_ARG, "vah_profile_read: NULL argument");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    memset(out, 0, sizeof *out);
    for (size_t k = 0; k < c->ev_kind.size(); ++k) {
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, c->ev[2 * k], c->ev[2 * k + 1]));
        out->avg_ms[c->ev_kind[k]] += ms;
        out->launches[c->ev_kind[k]]++;
    }
    for (int s = 0; s < 2; ++s)
        if (out->launches[s]) out->avg_ms[s] /= out->launches[s];
    return VAH_OK;
}

}  // extern "C"
