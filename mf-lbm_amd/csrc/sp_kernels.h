// Device code shared by the single-sample (sp_device.hip) and batched (sp_batch.hip) single-phase paths:
// the per-sample device view, the FP64/FP32 storage accessors, the per-node lattice update (singlephase_3D/
// 0.src/Kernel.F90:5-400), dense staging <-> compact storage, and the periodic self-exchange (Mpi.F90:73-448).
//
// Two storage modes share every kernel through the element type T of the population slices:
//   double: the populations themselves, the reference arithmetic bit for bit;
//   float : the deviations d = f - REST from the model's rest equilibrium (collide_sp.h), collided in FP32 or,
//           in the mixed mode, in FP64 (rounded to FP32 once per store).
//   Boundaries, the periodic exchange, the macroscopic fields and every reduction read and write f in FP64
//   through SPop (f = d + REST on load, d = f - REST rounded once on store); the exchange moves stored values.
#pragma once
#include <hip/hip_runtime.h>

#include "collide_sp.h"
#include "device_common.h"
#include "lattice.h"

namespace mflbm_sp {

using namespace mflbm_dev;

template <class T>
struct SDev {
    Layout L;
    // slice q: compact slots [0, nc), then the link slots of direction opp(q) (device_common.h)
    T* S[19];
    const int* map;   // padded offset -> compact index
    const int* cpos;  // compact index -> padded offset
    int nf;           // interior fluid nodes
    int nc;           // all compact nodes
    const unsigned long long* wmask;   // per wave and direction: fluid nodes with a link (device_common.h)
    const int* wbase;                  // per wave and direction: links before the wave
    const int8_t* walls;
    double* w_in;     // (nx+2)(ny+2), reference layout
    double* cf;       // convective-outlet history: 19 planes (nx+2)(ny+2)
};

// stored value <-> population f (FP64)
template <class T>
struct Store;
template <>
struct Store<double> {
    static __device__ __forceinline__ double get(double v, int) { return v; }
    static __device__ __forceinline__ double put(double f, int) { return f; }
};
template <>
struct Store<float> {
    static __device__ __forceinline__ double get(float v, int q) { return (double)v + rest_weight(q); }
    static __device__ __forceinline__ float put(double f, int q) { return (float)(f - rest_weight(q)); }
};

// population f of slot q, read and written in FP64 whatever the storage type
template <class T>
struct SPRef {
    T* p;
    int q;
    __device__ __forceinline__ operator double() const { return Store<T>::get(*p, q); }
    __device__ __forceinline__ SPRef& operator=(double f) {
        *p = Store<T>::put(f, q);
        return *this;
    }
    __device__ __forceinline__ SPRef& operator=(const SPRef& o) { return *this = (double)o; }
};

// population accessor by (direction, padded offset); a support node's slot q facing a fluid node is that
// node's link slot of direction opp(q)
template <class T>
struct SPop {
    const SDev<T>* d;
    __device__ __forceinline__ T& raw(int q, int o) const {
        const int c = d->map[o];
        if (c >= d->nf) {
            const int x = d->map[o + d->L.off(q)];
            if (x >= 0 && x < d->nf) return d->S[q][link_slot(d->nc, d->wbase, d->wmask, x, lat::OPP[q])];
        }
        return d->S[q][c];
    }
    __device__ __forceinline__ SPRef<T> operator()(int q, int o) const { return SPRef<T>{&raw(q, o), q}; }
};
template <class T>
__device__ __forceinline__ SPop<T> spop(const SDev<T>& d) { return SPop<T>{&d}; }
template <class T>
__device__ __forceinline__ int plane_at(const SDev<T>& d, int i, int j) { return i + j * (d.L.nx + 2); }

// ODD = false: kernel_even (Kernel.F90:206-400): read the swapped own slots, write the own slots.
// ODD = true : kernel_odd  (Kernel.F90:5-200): pull f_q(x - e_q), push to f_opp(q)(x + e_q).
// The update of fluid node m (compact index, m < nf).  m & 63 must be the node's lane in its wave and all
// lanes of a wave must belong to the same sample: the link masks are per wave of 64 consecutive nodes.
// T is the storage type, C the arithmetic of the collision (C = T, or FP32 storage with an FP64 collision).
template <bool ODD, class T, class C>
__device__ __forceinline__ void sp_update_node(const SDev<T>& d, const SPCollideParamsT<C>& P, int m) {
    T a[19];
    unsigned nb[19];   // odd: per direction r, the index of the slot (x + e_r, opp(r)) in slice opp(r)
    a[0] = slot_ld(d.S[0], m);
    if (ODD) {
        odd_neighbour_slots<true>(d.map, d.L, d.nc, d.wbase, d.wmask, m, d.cpos[m], nb);
#pragma unroll
        for (int q = 1; q < 19; ++q) a[q] = slot_ld(d.S[q], nb[lat::OPP[q]]);   // slot q of the upwind node x - e_q
    } else {
#pragma unroll
        for (int q = 1; q < 19; ++q) a[q] = slot_ld(d.S[lat::OPP[q]], m);
    }
    if (sizeof(C) == sizeof(T)) {
        collide_single<C>((C*)a, P);
    } else {   // FP32 storage collided in FP64: each deviation is rounded once, on store
        C w[19];
#pragma unroll
        for (int q = 0; q < 19; ++q) w[q] = (C)a[q];
        collide_single<C>(w, P);
#pragma unroll
        for (int q = 0; q < 19; ++q) a[q] = (T)w[q];
    }
    slot_st(d.S[0], m, a[0]);
#pragma unroll
    for (int q = 1; q < 19; ++q) {
        if (ODD) slot_st(d.S[lat::OPP[q]], nb[q], a[q]);   // slot opp(q) of x + e_q
        else slot_st(d.S[q], m, a[q]);
    }
}

// block order: plain for FP64 storage, XCD-aware for FP32 storage (each measured the faster at 256^3, by 2 %)
template <bool ODD, class T, class C>
__global__ void __launch_bounds__(BLOCK) k_sp_collide(SDev<T> d, SPCollideParamsT<C> P, int per) {
    const int m = (sizeof(T) == sizeof(double) ? (int)blockIdx.x : xcd_block(per)) * BLOCK + (int)threadIdx.x;
    if (m >= d.nf) return;
    sp_update_node<ODD, T, C>(d, P, m);
}

// dense staging (populations f, FP64) <-> compact storage
template <class T>
__global__ void k_sp_gather(SDev<T> d, const double* __restrict__ dense, long ps) {
    const long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (t >= d.nc) return;
    const int n = d.cpos[t];
    for (int q = 0; q < 19; ++q) d.S[q][t] = Store<T>::put(dense[q * ps + n], q);
}
template <class T>
__global__ void k_sp_scatter(SDev<T> d, double* __restrict__ dense, long ps) {
    const long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (t >= d.nc) return;
    const int n = d.cpos[t];
    for (int q = 0; q < 19; ++q) dense[q * ps + n] = Store<T>::get(d.S[q][t], q);
}
// link slots <-> dense staging (after gather / after scatter)
template <bool TO_DENSE, class T>
__global__ void k_sp_links(SDev<T> d, double* __restrict__ dense, long ps) {
    const int m = blockIdx.x * BLOCK + threadIdx.x;
    if (m >= d.nf) return;
    const int n = d.cpos[m];
    for (int r = 1; r < 19; ++r) {
        const int i = link_slot(d.nc, d.wbase, d.wmask, m, r);
        if (i < 0) continue;
        const long o = n + d.L.off(r);
        const int q = lat::OPP[r];
        if (TO_DENSE) dense[q * ps + o] = Store<T>::get(d.S[q][i], q);
        else d.S[q][i] = Store<T>::put(dense[q * ps + o], q);
    }
}

// Upload of a folded domain (layout.h, Fold): the exchanged slot (q, y) of each image takes the value of its halo
// slot (q, h), as the push exchange after the first (odd) step leaves it; the dense staging keeps the state as
// given.  (A checkpoint written after either exchange already holds equal values.)
template <class T>
__global__ void k_sp_fold_upload(SDev<T> d, Fold F, const double* __restrict__ dense, long ps) {
    fold_exchanged(d.L, F, (long)blockIdx.x * BLOCK + threadIdx.x, [&](int h, int y, int q) {
        if (d.map[y] < 0) return;
        spop(d)(q, y) = dense[(long)q * ps + h];
    });
}

// ------------------------------------------------------------------------------------------------------
// periodic wrap: the reference's MPI exchange of one rank with itself (Mpi.F90:73-448)
// ------------------------------------------------------------------------------------------------------
static __constant__ int SZP[5] = {5, 11, 12, 15, 16};
static __constant__ int SZM[5] = {6, 14, 13, 18, 17};
static __constant__ int SYP[5] = {3, 7, 8, 15, 17};
static __constant__ int SYM[5] = {4, 10, 9, 16, 18};

// the stored value moves unchanged (same direction on both sides)
template <class T>
__device__ __forceinline__ void cp(const SDev<T>& d, int q, int dst, int src) {
    const SPop<T> f = spop(d);
    f.raw(q, dst) = f.raw(q, src);
}
// after an even step (mpi_pdf_pull_async_pack / _sync_update, Mpi.F90:73-256)
template <class T>
__device__ __forceinline__ void sp_pull_z(const SDev<T>& d, int t) {
    const Layout L = d.L;
    if (t >= L.nx * L.ny) return;
    const int i = t % L.nx + 1, j = t / L.nx + 1;
    for (int m = 0; m < 5; ++m) {
        cp(d, SZP[m], L.at(i, j, 0), L.at(i, j, L.nz));
        cp(d, SZM[m], L.at(i, j, L.nz + 1), L.at(i, j, 1));
    }
}
template <class T>
__device__ __forceinline__ void sp_pull_y(const SDev<T>& d, int t) {
    const Layout L = d.L;
    if (t >= L.nx * L.nz) return;
    const int i = t % L.nx + 1, k = t / L.nx + 1;
    for (int m = 0; m < 5; ++m) {
        cp(d, SYP[m], L.at(i, 0, k), L.at(i, L.ny, k));
        cp(d, SYM[m], L.at(i, L.ny + 1, k), L.at(i, 1, k));
    }
}
template <class T>
__device__ __forceinline__ void sp_pull_edge(const SDev<T>& d, int t) {
    const Layout L = d.L;
    const int i = t + 1;
    if (i > L.nx) return;
    cp(d, 18, L.at(i, L.ny + 1, L.nz + 1), L.at(i, 1, 1));
    cp(d, 16, L.at(i, L.ny + 1, 0), L.at(i, 1, L.nz));
    cp(d, 17, L.at(i, 0, L.nz + 1), L.at(i, L.ny, 1));
    cp(d, 15, L.at(i, 0, 0), L.at(i, L.ny, L.nz));
}
// after an odd step (mpi_pdf_push_async_pack / _sync_update, Mpi.F90:264-448): faces, then x-edges
template <class T>
__device__ __forceinline__ void sp_push_z(const SDev<T>& d, int t) {
    const Layout L = d.L;
    if (t >= L.nx * L.ny) return;
    const int i = t % L.nx + 1, j = t / L.nx + 1;
    for (int m = 0; m < 5; ++m) {
        cp(d, SZM[m], L.at(i, j, 1), L.at(i, j, L.nz + 1));
        cp(d, SZP[m], L.at(i, j, L.nz), L.at(i, j, 0));
    }
}
template <class T>
__device__ __forceinline__ void sp_push_y(const SDev<T>& d, int t) {
    const Layout L = d.L;
    if (t >= L.nx * L.nz) return;
    const int i = t % L.nx + 1, k = t / L.nx + 1;
    for (int m = 0; m < 5; ++m) {
        cp(d, SYM[m], L.at(i, 1, k), L.at(i, L.ny + 1, k));
        cp(d, SYP[m], L.at(i, L.ny, k), L.at(i, 0, k));
    }
}
template <class T>
__device__ __forceinline__ void sp_push_edge(const SDev<T>& d, int t) {
    const Layout L = d.L;
    const int i = t + 1;
    if (i > L.nx) return;
    cp(d, 15, L.at(i, L.ny, L.nz), L.at(i, 0, 0));
    cp(d, 17, L.at(i, L.ny, 1), L.at(i, 0, L.nz + 1));
    cp(d, 16, L.at(i, 1, L.nz), L.at(i, L.ny + 1, 0));
    cp(d, 18, L.at(i, 1, 1), L.at(i, L.ny + 1, L.nz + 1));
}

// one thread per face node (edges: per x)
template <class T>
__global__ void k_sp_pull_z(SDev<T> d) {
    sp_pull_z(d, (int)(blockIdx.x * BLOCK + threadIdx.x));
}
template <class T>
__global__ void k_sp_pull_y(SDev<T> d) {
    sp_pull_y(d, (int)(blockIdx.x * BLOCK + threadIdx.x));
}
template <class T>
__global__ void k_sp_pull_edge(SDev<T> d) {
    sp_pull_edge(d, (int)(blockIdx.x * BLOCK + threadIdx.x));
}
template <class T>
__global__ void k_sp_push_z(SDev<T> d) {
    sp_push_z(d, (int)(blockIdx.x * BLOCK + threadIdx.x));
}
template <class T>
__global__ void k_sp_push_y(SDev<T> d) {
    sp_push_y(d, (int)(blockIdx.x * BLOCK + threadIdx.x));
}
template <class T>
__global__ void k_sp_push_edge(SDev<T> d) {
    sp_push_edge(d, (int)(blockIdx.x * BLOCK + threadIdx.x));
}

// compute_macro_vars (Misc.F90:368-423): rho = sum f, u = sum e f - F/2 at fluid nodes.  Solid nodes hold 0
// (the fields are cleared when the geometry is uploaded and nothing else writes them).
template <class T>
__global__ void k_sp_macro(SDev<T> d, double force_z, double* u, double* v, double* w, double* rho) {
    const int m = blockIdx.x * BLOCK + threadIdx.x;
    if (m >= d.nf) return;
    double f[19];
#pragma unroll
    for (int q = 0; q < 19; ++q) f[q] = Store<T>::get(d.S[q][m], q);
    const double fx = 0.0, fy = 0.0, fz = force_z;
    const int o = d.cpos[m];
    rho[o] = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8] + f[9] + f[10] + f[11] + f[12] + f[13] + f[14] +
             f[15] + f[16] + f[17] + f[18];
    u[o] = f[1] - f[2] + f[7] - f[8] + f[9] - f[10] + f[11] - f[12] + f[13] - f[14] - 0.5 * fx;
    v[o] = f[3] - f[4] + f[7] + f[8] - f[9] - f[10] + f[15] - f[16] + f[17] - f[18] - 0.5 * fy;
    w[o] = f[5] - f[6] + f[11] + f[12] - f[13] - f[14] + f[15] + f[16] - f[17] - f[18] - 0.5 * fz;
}


}  // namespace mflbm_sp
