// Geometry preprocessing on the GPU: the boundary-node lists and wall normals of geometry_preprocessing_new
// (Geometry_preprocessing.F90:9-512), the same lists mflbm_geometry_new builds on the host, bit for bit.
//
// Steps, each a kernel over the wall copy extended by 10 ghost layers (x fastest):
//   1. ghost layers: x clamped (never periodic), y and z clamped or wrapped by the periodic flags
//      (Geometry_preprocessing.F90:54-83); the smoothing field starts as the 0/1 wall value
//   2. boundary marks: solid with a fluid neighbour -> 2, fluid with a solid neighbour -> -1 (:121-143); the
//      class of a node never changes, so reading the unmarked copy equals the reference's in-place sweep
//   3. four passes of 27-point smoothing (:145-169), ping-pong between two buffers
//   4. list extraction in the reference's x-fastest order: flags, an exclusive scan for the output positions,
//      then one thread per listed node writes its entry (neighbour list and weight for solid nodes, the
//      normalised ISO8 gradient for fluid nodes, :227-383, 430-512)
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <new>
#include <vector>

#include "device_common.h"
#include "geometry.h"
#include "lattice.h"
#include "mflbm.h"

using namespace mflbm_dev;

namespace {

constexpr int GL = geo::GL;

struct Ext {
    int nx, ny, nz;
    long sx, sy;
    __host__ __device__ long at(int i, int j, int k) const {
        return (long)(i + GL - 1) + (long)(j + GL - 1) * sx + (long)(k + GL - 1) * sy;
    }
    __host__ __device__ long size() const { return sy * (nz + 2 * GL); }
};

// ghost-layer source of a coordinate: clamp to [1, n], or the periodic image (n >= GL)
__device__ __forceinline__ int src_coord(int v, int n, int per) {
    if (per) return v < 1 ? v + n : (v > n ? v - n : v);
    return v < 1 ? 1 : (v > n ? n : v);
}

__global__ void k_extend(Ext E, const int8_t* __restrict__ wg, int jper, int kper, int8_t* __restrict__ t,
                         double* __restrict__ s1, double* __restrict__ s2) {
    const long n = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (n >= E.size()) return;
    const int i = (int)(n % E.sx) + 1 - GL;
    const int j = (int)((n / E.sx) % (E.ny + 2 * GL)) + 1 - GL;
    const int k = (int)(n / E.sy) + 1 - GL;
    const int si = src_coord(i, E.nx, 0), sj = src_coord(j, E.ny, jper), sk = src_coord(k, E.nz, kper);
    const int8_t v = wg[(long)(si - 1) + (long)(sj - 1) * E.nx + (long)(sk - 1) * E.nx * E.ny];
    t[n] = v;
    s1[n] = (double)v;
    s2[n] = (double)v;
}

__device__ __forceinline__ bool inner(const Ext& E, int i, int j, int k) {
    return i >= 2 - GL && i <= E.nx + GL - 1 && j >= 2 - GL && j <= E.ny + GL - 1 && k >= 2 - GL && k <= E.nz + GL - 1;
}

__global__ void k_marks(Ext E, const int8_t* __restrict__ t, int8_t* __restrict__ t2) {
    const long n = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (n >= E.size()) return;
    const int i = (int)(n % E.sx) + 1 - GL;
    const int j = (int)((n / E.sx) % (E.ny + 2 * GL)) + 1 - GL;
    const int k = (int)(n / E.sy) + 1 - GL;
    int8_t c = t[n];
    if (inner(E, i, j, k)) {
        if (c == 1) {
            for (int q = 1; q <= 18; ++q)
                if (t[E.at(i + lat::EX[q], j + lat::EY[q], k + lat::EZ[q])] <= 0) {
                    c = 2;
                    break;
                }
        } else if (c == 0) {
            for (int q = 1; q <= 18; ++q)
                if (t[E.at(i + lat::EX[q], j + lat::EY[q], k + lat::EZ[q])] >= 1) {
                    c = -1;
                    break;
                }
        }
    }
    t2[n] = c;
}

__global__ void k_smooth(Ext E, const double* __restrict__ a, double* __restrict__ b) {
    const long n = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (n >= E.size()) return;
    const int i = (int)(n % E.sx) + 1 - GL;
    const int j = (int)((n / E.sx) % (E.ny + 2 * GL)) + 1 - GL;
    const int k = (int)(n / E.sy) + 1 - GL;
    if (!inner(E, i, j, k)) return;
    auto S = [&](int ii, int jj, int kk) { return a[E.at(ii, jj, kk)]; };
    b[n] = geo::smooth27(S, i, j, k);
}

// box [lo, hi]^3 (per axis relative to n) enumerated x fastest
struct Box {
    int lo;           // first coordinate on every axis
    int bx, by, bz;   // extents
    __device__ void ijk(long t, int& i, int& j, int& k) const {
        i = (int)(t % bx) + lo;
        j = (int)((t / bx) % by) + lo;
        k = (int)(t / ((long)bx * by)) + lo;
    }
    long size() const { return (long)bx * by * bz; }
};

__global__ void k_flags(Ext E, Box B, const int8_t* __restrict__ t2, int8_t want, int* __restrict__ flag) {
    const long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (t >= (long)B.bx * B.by * B.bz) return;
    int i, j, k;
    B.ijk(t, i, j, k);
    flag[t] = t2[E.at(i, j, k)] == want ? 1 : 0;
}

__global__ void k_solid_list(Ext E, Box B, const int8_t* __restrict__ t2, const int* __restrict__ flag,
                             const int* __restrict__ pos, int32_t* s_ijk, int32_t* s_nfl, int32_t* s_nbr, double* s_w) {
    const long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (t >= (long)B.bx * B.by * B.bz || !flag[t]) return;
    int i, j, k;
    B.ijk(t, i, j, k);
    const long p = pos[t];
    int cnt = 0;
    double w = 0.0;
    int32_t* nb = s_nbr + 18 * p;
    for (int q = 1; q <= 18; ++q)
        if (t2[E.at(i + lat::EX[q], j + lat::EY[q], k + lat::EZ[q])] <= 0) {
            w = w + lat::W[q];
            nb[cnt++] = q;
        }
    for (int q = cnt; q < 18; ++q) nb[q] = 0;
    s_ijk[3 * p] = i;
    s_ijk[3 * p + 1] = j;
    s_ijk[3 * p + 2] = k;
    s_nfl[p] = cnt;
    s_w[p] = w;
}

__global__ void k_fluid_list(Ext E, Box B, const double* __restrict__ sm, const int* __restrict__ flag,
                             const int* __restrict__ pos, double theta, int32_t* f_ijk, double* f_nw, double* f_theta) {
    const long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (t >= (long)B.bx * B.by * B.bz || !flag[t]) return;
    int i, j, k;
    B.ijk(t, i, j, k);
    const long p = pos[t];
    auto S = [&](int ii, int jj, int kk) { return sm[E.at(ii, jj, kk)]; };
    double nwx, nwy, nwz;
    geo::iso8_gradient(S, i, j, k, nwx, nwy, nwz);
    const double tmp = 1.0 / (sqrt(nwx * nwx + nwy * nwy + nwz * nwz) + lat::EPS);
    f_ijk[3 * p] = i;
    f_ijk[3 * p + 1] = j;
    f_ijk[3 * p + 2] = k;
    f_nw[3 * p] = nwx * tmp;
    f_nw[3 * p + 1] = nwy * tmp;
    f_nw[3 * p + 2] = nwz * tmp;
    f_theta[p] = theta;
}

struct Scratch {
    std::vector<void*> bufs;
    ~Scratch() {
        for (void* b : bufs)
            if (b) (void)hipFree(b);
    }
    template <class T>
    hipError_t alloc(T** p, size_t n) {
        hipError_t e = hipMalloc((void**)p, sizeof(T) * (n ? n : 1));
        if (e == hipSuccess) bufs.push_back((void*)*p);
        return e;
    }
};

#define GCHK(call)                                      \
    do {                                                \
        if ((call) != hipSuccess) return MFLBM_ERR_HIP; \
    } while (0)

// flags + exclusive scan over one box: positions of the listed nodes and their count
int scan_box(Scratch& S, const Ext& E, const Box& B, const int8_t* t2, int8_t want, int** flag, int** pos, long* count) {
    const long nb = B.size();
    GCHK(S.alloc(flag, nb));
    GCHK(S.alloc(pos, nb));
    hipLaunchKernelGGL(k_flags, dim3(nblocks(nb)), dim3(BLOCK), 0, 0, E, B, t2, want, *flag);
    GCHK(hipGetLastError());
    size_t tmp_bytes = 0;
    GCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, *flag, *pos, (int)nb));
    void* tmp = nullptr;
    GCHK(S.alloc((char**)&tmp, tmp_bytes));
    GCHK(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, *flag, *pos, (int)nb));
    int last_pos = 0, last_flag = 0;
    GCHK(hipMemcpy(&last_pos, *pos + nb - 1, sizeof(int), hipMemcpyDeviceToHost));
    GCHK(hipMemcpy(&last_flag, *flag + nb - 1, sizeof(int), hipMemcpyDeviceToHost));
    *count = (long)last_pos + last_flag;
    return MFLBM_OK;
}

}  // namespace

extern "C" int mflbm_geometry_new_device(const mflbm_params* p, const int8_t* walls_global, double theta,
                                         mflbm_geometry** out) {
    if (!p || !walls_global || !out || p->nx < 3 || p->ny < 3 || p->nz < 3) return MFLBM_ERR_ARG;
    *out = nullptr;
    if ((p->kper == 1 && p->nz < GL) || (p->jper == 1 && p->ny < GL)) return MFLBM_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MFLBM_ERR_NODEV;
    Ext E{p->nx, p->ny, p->nz, p->nx + 2 * GL, (long)(p->nx + 2 * GL) * (p->ny + 2 * GL)};
    const long ne = E.size();
    if ((long)p->nx * p->ny * p->nz >= (1L << 31) || ne >= (1L << 31)) return MFLBM_ERR_ARG;
    Scratch S;
    int8_t *wg = nullptr, *t = nullptr, *t2 = nullptr;
    double *a = nullptr, *b = nullptr;
    const size_t nw = (size_t)p->nx * p->ny * p->nz;
    GCHK(S.alloc(&wg, nw));
    GCHK(S.alloc(&t, ne));
    GCHK(S.alloc(&t2, ne));
    GCHK(S.alloc(&a, ne));
    GCHK(S.alloc(&b, ne));
    GCHK(hipMemcpy(wg, walls_global, nw, hipMemcpyHostToDevice));
    const unsigned g = nblocks(ne);
    hipLaunchKernelGGL(k_extend, dim3(g), dim3(BLOCK), 0, 0, E, (const int8_t*)wg, p->jper, p->kper, t, a, b);
    GCHK(hipGetLastError());
    hipLaunchKernelGGL(k_marks, dim3(g), dim3(BLOCK), 0, 0, E, (const int8_t*)t, t2);
    GCHK(hipGetLastError());
    for (int it = 0; it < 4; ++it) {
        const double* src = (it % 2 == 0) ? a : b;
        double* dst = (it % 2 == 0) ? b : a;
        hipLaunchKernelGGL(k_smooth, dim3(g), dim3(BLOCK), 0, 0, E, src, dst);
        GCHK(hipGetLastError());
    }
    const double* sm = a;   // after an even number of passes the result is back in a
    // solid list over [-2, n+3]^3, fluid list over [-1, n+2]^3 (Geometry_preprocessing.F90:430-445)
    const Box BS{-2, p->nx + 6, p->ny + 6, p->nz + 6};
    const Box BF{-1, p->nx + 4, p->ny + 4, p->nz + 4};
    int *fs = nullptr, *ps = nullptr, *ff = nullptr, *pf = nullptr;
    long ns = 0, nf = 0;
    int rc = scan_box(S, E, BS, t2, 2, &fs, &ps, &ns);
    if (rc) return rc;
    rc = scan_box(S, E, BF, t2, -1, &ff, &pf, &nf);
    if (rc) return rc;
    int32_t *s_ijk = nullptr, *s_nfl = nullptr, *s_nbr = nullptr, *f_ijk = nullptr;
    double *s_w = nullptr, *f_nw = nullptr, *f_th = nullptr;
    GCHK(S.alloc(&s_ijk, 3 * ns));
    GCHK(S.alloc(&s_nfl, ns));
    GCHK(S.alloc(&s_nbr, 18 * ns));
    GCHK(S.alloc(&s_w, ns));
    GCHK(S.alloc(&f_ijk, 3 * nf));
    GCHK(S.alloc(&f_nw, 3 * nf));
    GCHK(S.alloc(&f_th, nf));
    hipLaunchKernelGGL(k_solid_list, dim3(nblocks(BS.size())), dim3(BLOCK), 0, 0, E, BS, (const int8_t*)t2,
                       (const int*)fs, (const int*)ps, s_ijk, s_nfl, s_nbr, s_w);
    GCHK(hipGetLastError());
    hipLaunchKernelGGL(k_fluid_list, dim3(nblocks(BF.size())), dim3(BLOCK), 0, 0, E, BF, sm, (const int*)ff,
                       (const int*)pf, theta, f_ijk, f_nw, f_th);
    GCHK(hipGetLastError());
    GCHK(hipDeviceSynchronize());
    mflbm_geometry* gm = new (std::nothrow) mflbm_geometry();
    if (!gm) return MFLBM_ERR_NOMEM;
    gm->s_ijk.resize(3 * ns);
    gm->s_nfl.resize(ns);
    gm->s_nbr.resize(18 * ns);
    gm->s_w.resize(ns);
    gm->f_ijk.resize(3 * nf);
    gm->f_nw.resize(3 * nf);
    gm->f_theta.resize(nf);
    bool ok = true;
    auto down = [&](auto& v, const void* src) {
        if (!v.empty() && hipMemcpy(v.data(), src, v.size() * sizeof(v[0]), hipMemcpyDeviceToHost) != hipSuccess) ok = false;
    };
    down(gm->s_ijk, s_ijk);
    down(gm->s_nfl, s_nfl);
    down(gm->s_nbr, s_nbr);
    down(gm->s_w, s_w);
    down(gm->f_ijk, f_ijk);
    down(gm->f_nw, f_nw);
    down(gm->f_theta, f_th);
    if (!ok) {
        delete gm;
        return MFLBM_ERR_HIP;
    }
    *out = gm;
    return MFLBM_OK;
}
