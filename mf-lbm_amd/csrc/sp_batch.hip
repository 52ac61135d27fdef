// Batched single-phase permeability runs on one GPU (the mflbm_spb_* entry points of mflbm.h).
//
// Many samples (REV subvolumes of uneven size, each a complete single-phase problem of singlephase_3D/0.src:
// x walls, z periodic, body force along z) advance together.  Each sample keeps its own compact storage,
// dense map and link masks (device_common.h); the samples' fluid nodes are numbered into one wave-aligned
// list, so one launch of the lattice update covers every sample, with the wave's sample read from a table.
// The periodic exchange runs one launch per face kind for all samples (grid y = sample).
//
// Steady runs: a monitor period of ntime_monitor steps followed by the flow-rate reduction (FP64, per wave
// then per sample in a fixed order) and the reference's convergence test (Monitor.F90:145-156) is captured
// once as a hipGraph and replayed.  A sample that has converged is switched off on the device (its waves
// return at once) and keeps its state; the host waits once per period.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "batch_plan.h"
#include "device_common.h"
#include "lattice.h"
#include "mflbm.h"
#include "sp_kernels.h"

using namespace mflbm_dev;
using namespace mflbm_sp;

namespace {

constexpr int WAVE = 64;

// lattice update of every active sample; global wave w covers nodes [64 (w - wave_first[s]), +64) of sample s
template <bool ODD, class T, class C, int NT>
__global__ void __launch_bounds__(NT) k_spb_collide(const SDev<T>* __restrict__ devs, const int* __restrict__ wave_sample,
                                                    const int* __restrict__ wave_first, const int* __restrict__ active,
                                                    SPCollideParamsT<C> P, int nwaves, int per) {
    const int w = xcd_block(per) * (NT / WAVE) + (int)(threadIdx.x / WAVE);
    if (w >= nwaves) return;
    const int s = __builtin_amdgcn_readfirstlane(wave_sample[w]);
    if (!active[s]) return;
    const SDev<T>& d = devs[s];
    const int m = (w - wave_first[s]) * WAVE + (int)(threadIdx.x % WAVE);
    if (m >= d.nf) return;
    sp_update_node<ODD, T, C>(d, P, m);
}

// periodic self-exchange of every active sample periodic along y as well as z (the samples periodic along z only
// have it folded into their neighbour map, batch_plan.h); KIND 0..5 = pull z, pull y, pull edge, push z, push y,
// push edge (Mpi.F90:73-448)
template <class T, int KIND>
__global__ void k_spb_wrap(const SDev<T>* __restrict__ devs, const int* __restrict__ active,
                           const int* __restrict__ yper) {
    const int s = blockIdx.y;
    if (!active[s] || !yper[s]) return;
    const SDev<T>& d = devs[s];
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (KIND == 0) sp_pull_z(d, t);
    if (KIND == 1) sp_pull_y(d, t);
    if (KIND == 2) {
        if (t < d.L.nx) sp_pull_edge(d, t);
    }
    if (KIND == 3) sp_push_z(d, t);
    if (KIND == 4) sp_push_y(d, t);
    if (KIND == 5) {
        if (t < d.L.nx) sp_push_edge(d, t);
    }
}

// per-wave sum of w = sum_q e_z f_q - F/2 over the wave's fluid nodes (compute_macro_vars, Misc.F90:368-423;
// meaningful after even steps); butterfly reduction, the same order on every call
template <class T>
__global__ void __launch_bounds__(BLOCK) k_spb_flux(const SDev<T>* __restrict__ devs, const int* __restrict__ wave_sample,
                                                    const int* __restrict__ wave_first, const int* __restrict__ active,
                                                    double force_z, int nwaves, double* __restrict__ wave_sum) {
    const int w = (int)blockIdx.x * (BLOCK / WAVE) + (int)(threadIdx.x / WAVE);
    if (w >= nwaves) return;
    const int s = __builtin_amdgcn_readfirstlane(wave_sample[w]);
    if (!active[s]) return;
    const SDev<T>& d = devs[s];
    const int m = (w - wave_first[s]) * WAVE + (int)(threadIdx.x % WAVE);
    double v = 0.0;
    if (m < d.nf) {
        double f[19];
#pragma unroll
        for (int q = 0; q < 19; ++q) f[q] = Store<T>::get(d.S[q][m], q);
        v = f[5] - f[6] + f[11] + f[12] - f[13] - f[14] + f[15] + f[16] - f[17] - f[18] - 0.5 * force_z;
    }
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    if (threadIdx.x % WAVE == 0) wave_sum[w] = v;
}

struct Monitor {
    int* active;        // 1 while the sample runs
    int* conv_step;     // step of convergence (> 0), of failure (< 0), 0 while running
    double* q_prev;     // flow rate at the previous check
    double* q_last;
    int* period;        // monitor periods done
};

// one block per sample: flow rate Q = sum(w) / nz and the reference's test |Q - Q_prev| / (|Q| + eps) < tol
// at ntime > ntime_monitor (Monitor.F90:110-117,145-156); a NaN flow rate stops the sample as failed (:158-165)
__global__ void __launch_bounds__(BLOCK) k_spb_converge(const int* __restrict__ wave_first, const int* __restrict__ nwave,
                                                        const int* __restrict__ nzs, const double* __restrict__ wave_sum,
                                                        Monitor M, double tol, int ntime_monitor) {
    const int s = blockIdx.x;
    if (!M.active[s]) return;
    __shared__ double red[BLOCK];
    double v = 0.0;
    for (int i = threadIdx.x; i < nwave[s]; i += BLOCK) v += wave_sum[wave_first[s] + i];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int h = BLOCK / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double q = red[0] / (double)nzs[s];
        const int ntime = (M.period[0] + 1) * ntime_monitor;
        const double rel = fabs(q - M.q_prev[s]) / (fabs(q) + 1e-14);
        M.q_prev[s] = q;
        M.q_last[s] = q;
        if (q != q) {
            M.active[s] = 0;
            M.conv_step[s] = -ntime;
        } else if (ntime > ntime_monitor && rel < tol) {
            M.active[s] = 0;
            M.conv_step[s] = ntime;
        }
    }
}

// closes a period: counts it and the samples still running
__global__ void k_spb_tick(Monitor M, int nsamples, int* nactive) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    M.period[0] += 1;
    int n = 0;
    for (int s = 0; s < nsamples; ++s) n += M.active[s];
    nactive[0] = n;
}

template <class T>
__global__ void k_spb_fill(T* __restrict__ p, long n, T v) {
    const long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (t < n) p[t] = v;
}

__global__ void k_spb_set_int(int* __restrict__ p, int n, int v) {
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t < n) p[t] = v;
}

struct SampleHost {
    int nx = 0, ny = 0, nz = 0, jper = 0;
    std::vector<int8_t> walls_h2;
};

}  // namespace

struct mflbm_spb_ctx {
    mflbm_spb_params p{};
    SPCollideParams cp{};
    SPCollideParamsT<float> cp32{};
    bool fp32 = false;             // precision 1 or 2: FP32 deviation storage
    bool mixed = false;            // precision 2: collided in FP64
    int block = 256, tile = 32;
    std::vector<SampleHost> host;
    // per sample after commit
    std::vector<Layout> L;
    std::vector<int> nf, nc, wave_first_h, nwave_h;
    std::vector<long> elem_off;          // first element of the sample's 19 slices in `pop`
    std::vector<long> soff;              // 20 per sample: slice offsets inside the sample
    std::vector<SDev<double>> dev64;
    std::vector<SDev<float>> dev32;
    int nwaves = 0;
    long nfluid = 0;
    hipStream_t stream = nullptr;
    void* pop = nullptr;                 // all samples' population slices
    std::vector<void*> bufs;             // maps, cpos, link masks and bases of every sample
    void* devs = nullptr;                // SDev<T>[n]
    int* wave_sample = nullptr;
    int* wave_first = nullptr;
    int* nwave = nullptr;
    int* nzs = nullptr;
    int* yper = nullptr;                 // per sample: periodic along y too (exchange kernels, no fold)
    bool any_yper = false;
    double* wave_sum = nullptr;
    Monitor M{};
    int* mon_int = nullptr;              // active, conv_step, period, nactive
    double* mon_dbl = nullptr;           // q_prev, q_last
    int* nactive_h = nullptr;            // pinned
    double* staging = nullptr;           // dense upload/download buffer (19 arrays of the largest sample)
    long staging_len = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    bool committed = false;
    std::string err;
};

namespace {

template <class T>
const SDev<T>* devs_of(const mflbm_spb_ctx* c) {
    return (const SDev<T>*)c->devs;
}
template <class T>
const SDev<T>& sample_dev(const mflbm_spb_ctx* c, int s);
template <>
const SDev<double>& sample_dev<double>(const mflbm_spb_ctx* c, int s) { return c->dev64[s]; }
template <>
const SDev<float>& sample_dev<float>(const mflbm_spb_ctx* c, int s) { return c->dev32[s]; }
template <class T>
const SPCollideParamsT<T>& rates_of(const mflbm_spb_ctx* c);
template <>
const SPCollideParamsT<double>& rates_of<double>(const mflbm_spb_ctx* c) { return c->cp; }
template <>
const SPCollideParamsT<float>& rates_of<float>(const mflbm_spb_ctx* c) { return c->cp32; }

template <class T, class C, int NT>
int launch_collide_t(mflbm_spb_ctx* c, bool odd) {
    const long nb = ((long)c->nwaves * WAVE + NT - 1) / NT;
    const int per = (int)((nb + NXCD - 1) / NXCD);
    const unsigned grid = (unsigned)(NXCD * per);
    if (grid == 0) return MFLBM_OK;
    if (odd)
        hipLaunchKernelGGL((k_spb_collide<true, T, C, NT>), dim3(grid), dim3(NT), 0, c->stream, devs_of<T>(c),
                           c->wave_sample, c->wave_first, c->M.active, rates_of<C>(c), c->nwaves, per);
    else
        hipLaunchKernelGGL((k_spb_collide<false, T, C, NT>), dim3(grid), dim3(NT), 0, c->stream, devs_of<T>(c),
                           c->wave_sample, c->wave_first, c->M.active, rates_of<C>(c), c->nwaves, per);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(c, MFLBM_ERR_HIP, "launch k_spb_collide: %s", hipGetErrorString(e));
    return MFLBM_OK;
}

template <class T, int NT>
int launch_collide(mflbm_spb_ctx* c, bool odd) {
    return c->mixed ? launch_collide_t<T, double, NT>(c, odd) : launch_collide_t<T, T, NT>(c, odd);
}

template <class T, int KIND>
int launch_wrap(mflbm_spb_ctx* c, long nthreads) {
    const unsigned gx = nblocks(nthreads);
    const unsigned n = (unsigned)c->host.size();
    if (gx == 0 || n == 0) return MFLBM_OK;
    hipLaunchKernelGGL((k_spb_wrap<T, KIND>), dim3(gx, n), dim3(BLOCK), 0, c->stream, devs_of<T>(c), c->M.active, c->yper);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(c, MFLBM_ERR_HIP, "launch k_spb_wrap: %s", hipGetErrorString(e));
    return MFLBM_OK;
}

template <class T>
int step_t(mflbm_spb_ctx* c, int ntime) {
    const bool odd = (ntime % 2) != 0;
    int rc = MFLBM_OK;
    switch (c->block) {
        case 64: rc = launch_collide<T, 64>(c, odd); break;
        case 128: rc = launch_collide<T, 128>(c, odd); break;
        case 512: rc = launch_collide<T, 512>(c, odd); break;
        default: rc = launch_collide<T, 256>(c, odd); break;
    }
    if (rc || !c->any_yper) return rc;
    long fz = 0, fy = 0, fe = 0;
    for (const Layout& L : c->L) {
        fz = std::max(fz, (long)L.nx * L.ny);
        fy = std::max(fy, (long)L.nx * L.nz);
        fe = std::max(fe, (long)L.nx);
    }
    if (!odd) {
        if ((rc = launch_wrap<T, 0>(c, fz))) return rc;
        if ((rc = launch_wrap<T, 1>(c, fy))) return rc;
        if ((rc = launch_wrap<T, 2>(c, fe))) return rc;
    } else {
        if ((rc = launch_wrap<T, 3>(c, fz))) return rc;
        if ((rc = launch_wrap<T, 4>(c, fy))) return rc;
        if ((rc = launch_wrap<T, 5>(c, fe))) return rc;
    }
    return MFLBM_OK;
}

int step(mflbm_spb_ctx* c, int ntime) { return c->fp32 ? step_t<float>(c, ntime) : step_t<double>(c, ntime); }

template <class T>
int monitor_t(mflbm_spb_ctx* c) {
    const unsigned g = (unsigned)((c->nwaves + BLOCK / WAVE - 1) / (BLOCK / WAVE));
    if (g > 0) {
        hipLaunchKernelGGL(k_spb_flux<T>, dim3(g), dim3(BLOCK), 0, c->stream, devs_of<T>(c), c->wave_sample, c->wave_first,
                           c->M.active, c->p.force_z, c->nwaves, c->wave_sum);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return fail(c, MFLBM_ERR_HIP, "launch k_spb_flux: %s", hipGetErrorString(e));
    }
    LAUNCH(c, k_spb_converge, (unsigned)c->host.size(), c->wave_first, c->nwave, c->nzs, (const double*)c->wave_sum, c->M,
           c->p.tol, c->p.ntime_monitor);
    hipLaunchKernelGGL(k_spb_tick, dim3(1), dim3(64), 0, c->stream, c->M, (int)c->host.size(), c->mon_int + 3 * c->host.size());
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(c, MFLBM_ERR_HIP, "launch k_spb_tick: %s", hipGetErrorString(e));
    return MFLBM_OK;
}

// one monitor period: ntime_monitor steps from an odd step, then the check
int period(mflbm_spb_ctx* c) {
    for (int j = 0; j < c->p.ntime_monitor; ++j) {
        const int rc = step(c, 1 + j);
        if (rc) return rc;
    }
    return c->fp32 ? monitor_t<float>(c) : monitor_t<double>(c);
}

void free_all(mflbm_spb_ctx* c) {
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->graph_exec = nullptr;
    c->graph = nullptr;
    for (void* b : c->bufs)
        if (b) (void)hipFree(b);
    c->bufs.clear();
    void* own[] = {c->pop, c->devs, c->wave_sample, c->wave_first, c->nwave, c->nzs, c->yper, c->wave_sum,
                   c->mon_int, c->mon_dbl, c->staging};
    for (void* b : own)
        if (b) (void)hipFree(b);
    if (c->nactive_h) (void)hipHostFree(c->nactive_h);
    c->pop = c->devs = nullptr;
    c->wave_sample = c->wave_first = c->nwave = c->nzs = c->yper = c->mon_int = c->nactive_h = nullptr;
    c->wave_sum = c->mon_dbl = c->staging = nullptr;
}

template <class T>
int init_state_t(mflbm_spb_ctx* c) {
    // fluid at rest with density 1 at every slot, link slots included: f_q = w_q as initialization_new with
    // u = 0 (Initialization.F90:207-307), or with rest_state = 1 the collision's own rest state REST_q
    for (size_t s = 0; s < c->host.size(); ++s)
        for (int q = 0; q < 19; ++q) {
            const long n = c->soff[20 * s + q + 1] - c->soff[20 * s + q];
            const double f0 = c->p.rest_state ? rest_weight(q) : lat::W[q];
            const T v = (T)(c->fp32 ? f0 - rest_weight(q) : f0);
            T* p = (T*)c->pop + c->elem_off[s] + c->soff[20 * s + q];
            LAUNCH(c, k_spb_fill<T>, nblocks(n), p, n, v);
        }
    return MFLBM_OK;
}

int reset_monitor(mflbm_spb_ctx* c) {
    const int n = (int)c->host.size();
    LAUNCH(c, k_spb_set_int, nblocks(n), c->M.active, n, 1);
    LAUNCH(c, k_spb_set_int, nblocks(n), c->M.conv_step, n, 0);
    LAUNCH(c, k_spb_set_int, 1u, c->M.period, 1, 0);
    LAUNCH(c, k_spb_fill<double>, nblocks(2L * n), c->mon_dbl, 2L * n, 0.0);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MFLBM_OK;
}

bool sample_ok(const mflbm_spb_ctx* c, int s) { return c->committed && s >= 0 && s < (int)c->host.size(); }

}  // namespace

extern "C" const char* mflbm_spb_last_error(const mflbm_spb_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

extern "C" int mflbm_spb_create(const mflbm_spb_params* prm, mflbm_spb_ctx** out) {
    if (!out) return fail(nullptr, MFLBM_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!prm) return fail(nullptr, MFLBM_ERR_ARG, "params is NULL");
    if (prm->precision < 0 || prm->precision > 2)
        return fail(nullptr, MFLBM_ERR_ARG, "precision must be 0 (FP64), 1 (FP32) or 2 (FP32 storage, FP64 collision), got %d",
                    prm->precision);
    if (!(prm->la_nu > 0.0)) return fail(nullptr, MFLBM_ERR_ARG, "viscosity must be positive");
    if (prm->ntime_monitor < 2 || prm->ntime_monitor % 2)
        return fail(nullptr, MFLBM_ERR_ARG, "ntime_monitor must be even and positive (the flow rate is read after even steps)");
    if (prm->block != 0 && prm->block != 64 && prm->block != 128 && prm->block != 256 && prm->block != 512)
        return fail(nullptr, MFLBM_ERR_ARG, "block must be 0, 64, 128, 256 or 512");
    if (prm->tile_rows < 0) return fail(nullptr, MFLBM_ERR_ARG, "tile_rows must be >= 0");
    if (prm->rest_state != 0 && prm->rest_state != 1) return fail(nullptr, MFLBM_ERR_ARG, "rest_state must be 0 or 1");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(nullptr, MFLBM_ERR_NODEV, "no HIP device visible");
    mflbm_spb_ctx* c = new (std::nothrow) mflbm_spb_ctx();
    if (!c) return fail(nullptr, MFLBM_ERR_NOMEM, "context allocation failed");
    c->p = *prm;
    c->cp = sp_collide_params(prm->la_nu, prm->mrt_preset, prm->force_z);
    c->cp32 = to_float(c->cp);
    c->fp32 = prm->precision != 0;
    c->mixed = prm->precision == 2;
    c->block = prm->block ? prm->block : 256;
    c->tile = prm->tile_rows ? prm->tile_rows : tile_rows();
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(nullptr, MFLBM_ERR_HIP, "hipStreamCreate failed");
    }
    *out = c;
    return MFLBM_OK;
}

extern "C" void mflbm_spb_destroy(mflbm_spb_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_all(c);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" int mflbm_spb_add_sample(mflbm_spb_ctx* c, int32_t nx, int32_t ny, int32_t nz, int32_t jper,
                                    const int8_t* walls_h2) {
    if (!c) return MFLBM_ERR_ARG;
    if (c->committed) return fail(c, MFLBM_ERR_STATE, "samples are added before mflbm_spb_commit");
    if (nx < 3 || ny < 3 || nz < 3) return fail(c, MFLBM_ERR_ARG, "sample must be at least 3^3");
    if (jper != 0 && jper != 1) return fail(c, MFLBM_ERR_ARG, "jper must be 0 or 1");
    if (!walls_h2) return fail(c, MFLBM_ERR_ARG, "walls is NULL");
    SampleHost h;
    h.nx = nx;
    h.ny = ny;
    h.nz = nz;
    h.jper = jper;
    h.walls_h2.assign(walls_h2, walls_h2 + (size_t)(nx + 4) * (ny + 4) * (nz + 4));
    c->host.push_back(std::move(h));
    return MFLBM_OK;
}

extern "C" int mflbm_spb_commit(mflbm_spb_ctx* c) {
    if (!c) return MFLBM_ERR_ARG;
    if (c->committed) return fail(c, MFLBM_ERR_STATE, "already committed");
    const int n = (int)c->host.size();
    if (n == 0) return fail(c, MFLBM_ERR_ARG, "no samples");
    const long esz = c->fp32 ? (long)sizeof(float) : (long)sizeof(double);
    std::vector<BatchSampleIn> in(n);
    for (int s = 0; s < n; ++s) {
        const SampleHost& h = c->host[s];
        in[s] = BatchSampleIn{h.nx, h.ny, h.nz, h.jper, h.walls_h2.data()};
    }
    BatchPlan P;
    {
        const std::string why = plan_batch(in, c->tile, esz, P);
        if (!why.empty()) return fail(c, MFLBM_ERR_ARG, "%s", why.c_str());
    }
    const std::vector<CompactSet>& cs = P.cs;
    const std::vector<LinkSet>& ls = P.ls;
    c->L = P.L;
    c->nf = P.nf;
    c->nc = P.nc;
    c->wave_first_h = P.wave_first;
    c->nwave_h = P.nwave;
    c->elem_off = P.elem_off;
    c->soff = P.soff;
    c->nfluid = P.nfluid;
    c->nwaves = P.nwaves;
    const long elems = P.elems, maxdense = P.maxdense;
    const int waves = P.nwaves;
    // device buffers
    HIPCHK(c, hipMalloc(&c->pop, (size_t)(esz * std::max(elems, 1L))));
    std::vector<SDev<double>> d64(n);
    std::vector<SDev<float>> d32(n);
    for (int s = 0; s < n; ++s) {
        const Layout& L = c->L[s];
        int *map = nullptr, *cpos = nullptr, *wbase = nullptr;
        unsigned long long* wmask = nullptr;
        HIPCHK(c, hipMalloc(&map, sizeof(int) * L.total));
        c->bufs.push_back(map);
        HIPCHK(c, hipMalloc(&cpos, sizeof(int) * ((size_t)c->nc[s] + 1)));
        c->bufs.push_back(cpos);
        HIPCHK(c, hipMalloc(&wmask, sizeof(unsigned long long) * ls[s].wmask.size()));
        c->bufs.push_back(wmask);
        HIPCHK(c, hipMalloc(&wbase, sizeof(int) * ls[s].wbase.size()));
        c->bufs.push_back(wbase);
        HIPCHK(c, hipMemcpy(map, cs[s].cidx.data(), sizeof(int) * L.total, hipMemcpyHostToDevice));
        if (c->nc[s] > 0) HIPCHK(c, hipMemcpy(cpos, cs[s].cpos.data(), sizeof(int) * c->nc[s], hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(wmask, ls[s].wmask.data(), sizeof(unsigned long long) * ls[s].wmask.size(), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(wbase, ls[s].wbase.data(), sizeof(int) * ls[s].wbase.size(), hipMemcpyHostToDevice));
        const long* so = &c->soff[20L * s];
        d64[s].L = L;
        d64[s].map = map;
        d64[s].cpos = cpos;
        d64[s].nf = c->nf[s];
        d64[s].nc = c->nc[s];
        d64[s].wmask = wmask;
        d64[s].wbase = wbase;
        d64[s].walls = nullptr;
        d64[s].w_in = nullptr;
        d64[s].cf = nullptr;
        d32[s].L = L;
        d32[s].map = map;
        d32[s].cpos = cpos;
        d32[s].nf = c->nf[s];
        d32[s].nc = c->nc[s];
        d32[s].wmask = wmask;
        d32[s].wbase = wbase;
        d32[s].walls = nullptr;
        d32[s].w_in = nullptr;
        d32[s].cf = nullptr;
        for (int q = 0; q < 19; ++q) {
            d64[s].S[q] = c->fp32 ? nullptr : (double*)c->pop + c->elem_off[s] + so[q];
            d32[s].S[q] = c->fp32 ? (float*)c->pop + c->elem_off[s] + so[q] : nullptr;
        }
    }
    c->dev64 = d64;
    c->dev32 = d32;
    const size_t dsz = c->fp32 ? sizeof(SDev<float>) : sizeof(SDev<double>);
    HIPCHK(c, hipMalloc(&c->devs, dsz * n));
    if (c->fp32) HIPCHK(c, hipMemcpy(c->devs, d32.data(), dsz * n, hipMemcpyHostToDevice));
    else HIPCHK(c, hipMemcpy(c->devs, d64.data(), dsz * n, hipMemcpyHostToDevice));
    const std::vector<int>& ws = P.wave_sample;
    std::vector<int> nzs(n), yp(n);
    for (int s = 0; s < n; ++s) {
        nzs[s] = c->host[s].nz;
        yp[s] = c->host[s].jper;
        c->any_yper = c->any_yper || yp[s] == 1;
    }
    HIPCHK(c, hipMalloc(&c->wave_sample, sizeof(int) * ws.size()));
    HIPCHK(c, hipMalloc(&c->wave_first, sizeof(int) * n));
    HIPCHK(c, hipMalloc(&c->nwave, sizeof(int) * n));
    HIPCHK(c, hipMalloc(&c->nzs, sizeof(int) * n));
    HIPCHK(c, hipMalloc(&c->yper, sizeof(int) * n));
    HIPCHK(c, hipMalloc(&c->wave_sum, sizeof(double) * ws.size()));
    HIPCHK(c, hipMalloc(&c->mon_int, sizeof(int) * (3 * (size_t)n + 2)));
    HIPCHK(c, hipMalloc(&c->mon_dbl, sizeof(double) * 2 * (size_t)n));
    HIPCHK(c, hipHostMalloc((void**)&c->nactive_h, sizeof(int)));
    HIPCHK(c, hipMemcpy(c->wave_sample, ws.data(), sizeof(int) * ws.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->wave_first, c->wave_first_h.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->nwave, c->nwave_h.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->nzs, nzs.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->yper, yp.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    c->M.active = c->mon_int;
    c->M.conv_step = c->mon_int + n;
    c->M.period = c->mon_int + 2 * n;
    c->M.q_prev = c->mon_dbl;
    c->M.q_last = c->mon_dbl + n;
    c->staging_len = maxdense;
    c->committed = true;
    int rc = c->fp32 ? init_state_t<float>(c) : init_state_t<double>(c);
    if (rc) return rc;
    rc = reset_monitor(c);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MFLBM_OK;
}

extern "C" int mflbm_spb_info(mflbm_spb_ctx* c, int32_t* nsamples, int64_t* fluid_nodes, int64_t* pop_bytes) {
    if (!c) return MFLBM_ERR_ARG;
    if (!c->committed) return fail(c, MFLBM_ERR_STATE, "not committed");
    if (nsamples) *nsamples = (int32_t)c->host.size();
    if (fluid_nodes) *fluid_nodes = c->nfluid;
    if (pop_bytes) {
        long e = 0;
        for (size_t s = 0; s < c->host.size(); ++s) e += c->soff[20 * s + 19];
        *pop_bytes = e * (c->fp32 ? 4 : 8);
    }
    return MFLBM_OK;
}

extern "C" int mflbm_spb_sample_fluid(mflbm_spb_ctx* c, int32_t sample, int64_t* fluid_nodes) {
    if (!c) return MFLBM_ERR_ARG;
    if (!sample_ok(c, sample)) return fail(c, MFLBM_ERR_ARG, "no sample %d (or not committed)", sample);
    if (fluid_nodes) *fluid_nodes = c->nf[sample];
    return MFLBM_OK;
}

namespace {
int ensure_staging(mflbm_spb_ctx* c) {
    if (c->staging) return MFLBM_OK;
    HIPCHK(c, hipMalloc(&c->staging, sizeof(double) * 19 * (size_t)c->staging_len));
    return MFLBM_OK;
}
template <class T>
int gather_sample(mflbm_spb_ctx* c, int s) {
    const SDev<T>& d = sample_dev<T>(c, s);
    const long ps = c->L[s].total;
    LAUNCH(c, k_sp_gather<T>, nblocks(d.nc), d, (const double*)c->staging, ps);
    LAUNCH(c, (k_sp_links<false, T>), nblocks(d.nf), d, c->staging, ps);
    const Fold F = batch_fold(c->host[s].jper);
    if (F.y || F.z) LAUNCH(c, k_sp_fold_upload<T>, fold_blocks(c->L[s]), d, F, (const double*)c->staging, ps);
    return MFLBM_OK;
}
template <class T>
int scatter_sample(mflbm_spb_ctx* c, int s) {
    const SDev<T>& d = sample_dev<T>(c, s);
    const long ps = c->L[s].total;
    LAUNCH(c, k_sp_scatter<T>, nblocks(d.nc), d, c->staging, ps);
    LAUNCH(c, (k_sp_links<true, T>), nblocks(d.nf), d, c->staging, ps);
    const Fold F = batch_fold(c->host[s].jper);
    if (F.y || F.z) LAUNCH(c, k_fold_dense, fold_blocks(c->L[s]), c->L[s], F, d.map, d.nf, c->staging, ps, 1);
    return MFLBM_OK;
}
}  // namespace

extern "C" int mflbm_spb_upload_state(mflbm_spb_ctx* c, int32_t sample, const double* const* f) {
    if (!c) return MFLBM_ERR_ARG;
    if (!sample_ok(c, sample)) return fail(c, MFLBM_ERR_ARG, "no sample %d (or not committed)", sample);
    if (!f) return fail(c, MFLBM_ERR_ARG, "missing population arrays");
    for (int q = 0; q < 19; ++q)
        if (!f[q]) return fail(c, MFLBM_ERR_ARG, "population pointer %d is NULL", q);
    int rc = ensure_staging(c);
    if (rc) return rc;
    const Layout& L = c->L[sample];
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<double> buf(L.total, 0.0);
    for (int q = 0; q < 19; ++q) {
        pack_host(L, 1, f[q], buf.data());
        HIPCHK(c, hipMemcpy(c->staging + q * L.total, buf.data(), sizeof(double) * L.total, hipMemcpyHostToDevice));
    }
    rc = c->fp32 ? gather_sample<float>(c, sample) : gather_sample<double>(c, sample);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MFLBM_OK;
}

extern "C" int mflbm_spb_download_state(mflbm_spb_ctx* c, int32_t sample, double* const* f) {
    if (!c) return MFLBM_ERR_ARG;
    if (!sample_ok(c, sample)) return fail(c, MFLBM_ERR_ARG, "no sample %d (or not committed)", sample);
    if (!f) return fail(c, MFLBM_ERR_ARG, "missing population arrays");
    int rc = ensure_staging(c);
    if (rc) return rc;
    const Layout& L = c->L[sample];
    // slots outside the compact storage come back as the initial state (fluid at rest)
    for (int q = 0; q < 19; ++q)
        LAUNCH(c, k_spb_fill<double>, nblocks(L.total), c->staging + q * L.total, L.total,
               c->p.rest_state ? rest_weight(q) : lat::W[q]);
    rc = c->fp32 ? scatter_sample<float>(c, sample) : scatter_sample<double>(c, sample);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<double> buf(L.total);
    for (int q = 0; q < 19; ++q) {
        if (!f[q]) continue;
        HIPCHK(c, hipMemcpy(buf.data(), c->staging + q * L.total, sizeof(double) * L.total, hipMemcpyDeviceToHost));
        unpack_host(L, 1, buf.data(), f[q]);
    }
    return MFLBM_OK;
}

// macroscopic fields of one sample (compute_macro_vars, Misc.F90:368-423; meaningful after even steps) into u,
// v, w, rho (each may be NULL) with the halo-1 layout of mflbm_sp_download_macro; solid nodes hold 0
extern "C" int mflbm_spb_download_macro(mflbm_spb_ctx* c, int32_t sample, double* u, double* v, double* w,
                                        double* rho) {
    if (!c) return MFLBM_ERR_ARG;
    if (!sample_ok(c, sample)) return fail(c, MFLBM_ERR_ARG, "no sample %d (or not committed)", sample);
    int rc = ensure_staging(c);
    if (rc) return rc;
    const Layout& L = c->L[sample];
    const long ps = L.total;
    double* st = c->staging;
    HIPCHK(c, hipMemsetAsync(st, 0, sizeof(double) * 4 * ps, c->stream));
    if (c->fp32) {
        const SDev<float>& d = c->dev32[sample];
        LAUNCH(c, k_sp_macro<float>, nblocks(d.nf), d, c->p.force_z, st, st + ps, st + 2 * ps, st + 3 * ps);
    } else {
        const SDev<double>& d = c->dev64[sample];
        LAUNCH(c, k_sp_macro<double>, nblocks(d.nf), d, c->p.force_z, st, st + ps, st + 2 * ps, st + 3 * ps);
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<double> buf(L.total);
    double* dst[4] = {u, v, w, rho};
    for (int m = 0; m < 4; ++m) {
        if (!dst[m]) continue;
        HIPCHK(c, hipMemcpy(buf.data(), st + m * ps, sizeof(double) * L.total, hipMemcpyDeviceToHost));
        unpack_host(L, 1, buf.data(), dst[m]);
    }
    return MFLBM_OK;
}

extern "C" int mflbm_spb_run(mflbm_spb_ctx* c, int32_t ntime0, int32_t nsteps) {
    if (!c) return MFLBM_ERR_ARG;
    if (!c->committed) return fail(c, MFLBM_ERR_STATE, "not committed");
    if (nsteps < 0) return fail(c, MFLBM_ERR_ARG, "nsteps < 0");
    for (int s = 0; s < nsteps; ++s) {
        const int rc = step(c, ntime0 + s);
        if (rc) return rc;
    }
    return MFLBM_OK;
}

extern "C" int mflbm_spb_run_steady(mflbm_spb_ctx* c, int32_t max_periods, int32_t* nactive) {
    if (!c) return MFLBM_ERR_ARG;
    if (!c->committed) return fail(c, MFLBM_ERR_STATE, "not committed");
    if (max_periods < 0) return fail(c, MFLBM_ERR_ARG, "max_periods < 0");
    if (!c->graph_exec) {
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int rc = period(c);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(c, MFLBM_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
        c->graph = g;
        HIPCHK(c, hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0));
    }
    int left = 1;
    for (int k = 0; k < max_periods; ++k) {
        HIPCHK(c, hipGraphLaunch(c->graph_exec, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->nactive_h, c->mon_int + 3 * c->host.size(), sizeof(int), hipMemcpyDeviceToHost,
                                 c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        left = *c->nactive_h;
        if (left == 0) break;
    }
    if (nactive) *nactive = left;
    return MFLBM_OK;
}

extern "C" int mflbm_spb_results(mflbm_spb_ctx* c, int32_t* conv_step, double* flowrate, int32_t* periods) {
    if (!c) return MFLBM_ERR_ARG;
    if (!c->committed) return fail(c, MFLBM_ERR_STATE, "not committed");
    const size_t n = c->host.size();
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (conv_step) HIPCHK(c, hipMemcpy(conv_step, c->M.conv_step, sizeof(int) * n, hipMemcpyDeviceToHost));
    if (flowrate) HIPCHK(c, hipMemcpy(flowrate, c->M.q_last, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (periods) HIPCHK(c, hipMemcpy(periods, c->M.period, sizeof(int), hipMemcpyDeviceToHost));
    return MFLBM_OK;
}

extern "C" int mflbm_spb_reset_monitor(mflbm_spb_ctx* c) {
    if (!c) return MFLBM_ERR_ARG;
    if (!c->committed) return fail(c, MFLBM_ERR_STATE, "not committed");
    return reset_monitor(c);
}

extern "C" int mflbm_spb_synchronize(mflbm_spb_ctx* c) {
    if (!c) return MFLBM_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MFLBM_OK;
}
