// Pieces shared by the two-phase and the single-phase gfx950 paths: the padded dense layout, the compact
// node set the populations are stored on, the XCD-aware block order, host-side halo packing and the error
// plumbing of the C ABI.
//
// Dense layout (every dense field: phase field, normals, walls, macroscopic fields):
//
//   offset(i, j, k) = (i + XO) + (j + YO) * px + (k + ZO) * px * py      (1-based interior indices)
//
// with XO = 15 so that i = 1 sits on a 128-byte boundary, px = round_up(nx + 20, 16), py = ny + 8,
// pz = nz + 8 (at least the 4-layer halo of the phase field in every direction).
//
// Compact population storage: one slot per interior fluid node (compact indices [0, nf), x-fastest order)
// followed by the "support" nodes: every non-fluid node a fluid node reaches in one lattice link (solid
// neighbours, whose slots carry the wall reflection of the AA update, and halo nodes) and every node of the
// planes the boundary conditions and the periodic exchange address.  A dense int32 map gives the compact
// index of a padded offset (-1 where nothing is stored).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdlib>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "lattice.h"
#include "layout.h"
#include "mflbm.h"

namespace mflbm_dev {

constexpr int BLOCK = 256;
constexpr int NXCD = 8;   // gfx950: 8 accelerator complex dies, work-groups dispatched round-robin

// Logical block index with each XCD taking one contiguous range of blocks: node lists are in x-fastest
// order, so a die then works on a z-slab and stencil neighbours stay in its own L2.  The grid is launched
// as NXCD * per blocks.  (The lattice updates use the plain block order instead: with every die streaming its own
// distant range, they ran 2 % slower at 256^3, the populations having no reuse to gain from the L2.)
__device__ __forceinline__ int xcd_block(int per) { return (int)(blockIdx.x % NXCD) * per + (int)(blockIdx.x / NXCD); }

inline unsigned nblocks(long n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }
// grid of a list kernel that uses xcd_block(): per = blocks per die
inline int xcd_per(long n) { return (int)((nblocks(n) + NXCD - 1) / NXCD); }

// link slot of fluid node x (compact index) in direction r, or -1 when x + e_r is a fluid node
__device__ __forceinline__ int link_slot(int nc, const int* wbase, const unsigned long long* wmask, int x, int r) {
    const int e = 18 * (x >> 6) + r - 1;
    const unsigned long long wm = wmask[e];
    const unsigned l = (unsigned)x & 63u;
    if (!((wm >> l) & 1ull)) return -1;
    return nc + wbase[e] + __popcll(wm & ((1ull << l) - 1ull));
}

// The odd step's neighbour slots of fluid node m (compact index; m & 63 its lane, every lane of the wave in one link
// mask word) at dense offset n: per direction r, the index of the slot (x + e_r, opp(r)) in slice opp(r), which is
// the link slot of (m, r) where the wave's link mask has m's bit, else the compact index of the fluid node x + e_r.
// The wave-uniform masks and bases are scalar loads issued together; x + e_r lies inside the padded grid, so the
// 18 dense map entries are loaded unconditionally, all in flight at once, and selected afterwards (a conditional
// load per direction compiles to a branch that waits for every earlier load).  ROWS: the entries of a row of three
// directions (ex = -1, 0, +1 at one (ey, ez)) are adjacent and taken by one 12-byte load (10 load instructions
// instead of 18): 1.6-2.1 % faster single-phase steps at 256^3, 0.4 % slower for the two-phase update, whose
// curvature stencil keeps more loads in flight.
template <bool ROWS>
__device__ __forceinline__ void odd_neighbour_slots(const int* __restrict__ map, const Layout& L, int nc,
                                                    const int* __restrict__ wbase,
                                                    const unsigned long long* __restrict__ wmask, int m, int n,
                                                    unsigned nb[19]) {
    const int e0 = 18 * __builtin_amdgcn_readfirstlane(m >> 6) - 1;
    const unsigned lane = (unsigned)m & 63u;
    const unsigned long long below = (1ull << lane) - 1ull;
    unsigned mv[19];
    if constexpr (ROWS) {   // rows (ey, ez) = (0, 0), (1, 0), (-1, 0), (0, 1), (0, -1); the four directions with ey
                            // and ez both nonzero one by one (4-byte aligned 12-byte loads)
        struct __attribute__((aligned(4))) Row {
            int v[3];
        };
        Row row[5];
        const int rb[5] = {n - 1, n + L.px - 1, n - L.px - 1, n + L.pxy - 1, n - L.pxy - 1};
#pragma unroll
        for (int k = 0; k < 5; ++k) row[k] = *(const Row*)(map + rb[k]);
#pragma unroll
        for (int r = 1; r < 19; ++r) {
            const int ey = lat::EY[r], ez = lat::EZ[r];
            const int k = ey == 1 ? 1 : ey == -1 ? 2 : ez == 1 ? 3 : ez == -1 ? 4 : 0;
            mv[r] = (unsigned)((ey != 0 && ez != 0) ? map[n + L.off(r)] : row[k].v[lat::EX[r] + 1]);
        }
    } else {
#pragma unroll
        for (int r = 1; r < 19; ++r) mv[r] = (unsigned)map[n + L.off(r)];
    }
#pragma unroll
    for (int r = 1; r < 19; ++r) {
        const unsigned long long wm = wmask[e0 + r];
        const unsigned li = (unsigned)(nc + wbase[e0 + r]) + (unsigned)__popcll(wm & below);
        nb[r] = ((wm >> lane) & 1ull) ? li : mv[r];
    }
}

// Slot access with a scalar slice base and a 32-bit per-lane byte offset (one global load/store with an SGPR
// base instead of a 64-bit vector address per slot); slices are kept below 2^32 bytes.
template <class T>
__device__ __forceinline__ T slot_ld(const T* slice, unsigned i) {
    return *(const T*)((const char*)slice + i * (unsigned)sizeof(T));
}
template <class T>
__device__ __forceinline__ void slot_st(T* slice, unsigned i, T v) {
    *(T*)((char*)slice + i * (unsigned)sizeof(T)) = v;
}

// Exchanged halo slots of a folded domain (layout.h, Fold): the halo slot (q, h) of a direction q that points
// into the domain along h's halo axes, and the slot (q, y) of h's image.  The reference's self-exchange leaves them
// equal after every step.
//   fold_exchanged: calls f(h, y, q) for each such pair of the one-layer halo node (i, j, k) (x interior).
//   k_fold_dense (downloads): the compact storage kept the value at y, except where y is not fluid and the
//     interior node x = h + e_q is: then it is x's link slot of direction opp(q), which the link download wrote
//     at h, and y takes it from there; and where y holds no slot at all (no step touches the pair, and the first
//     push exchange left (q, y) equal to the (q, h) the staging still holds).  One thread per node of the
//     one-layer halo box; `ncol` population sets of 19 arrays each in the dense staging.
template <class Fn>
__device__ __forceinline__ void fold_exchanged(const Layout& L, const Fold& F, long t, Fn&& f) {
    const long ny2 = L.ny + 2, nz2 = L.nz + 2;
    if (t >= (long)L.nx * ny2 * nz2) return;
    const int i = (int)(t % L.nx) + 1;
    const int j = (int)((t / L.nx) % ny2);
    const int k = (int)(t / ((long)L.nx * ny2));
    int y;
    if (!fold_image(L, F, i, j, k, y)) return;
    const int h = L.at(i, j, k);
    const bool gy = j == 0 || j == L.ny + 1, gz = k == 0 || k == L.nz + 1;
    const int ey = j == 0 ? 1 : -1, ez = k == 0 ? 1 : -1;
#pragma unroll
    for (int q = 1; q < 19; ++q) {
        if ((gy && lat::EY[q] != ey) || (gz && lat::EZ[q] != ez)) continue;
        f(h, y, q);
    }
}
inline unsigned fold_blocks(const Layout& L) { return nblocks((long)L.nx * (L.ny + 2) * (L.nz + 2)); }
static __global__ void k_fold_dense(Layout L, Fold F, const int* __restrict__ map, int nf, double* __restrict__ dense,
                                    long ps, int ncol) {
    fold_exchanged(L, F, (long)blockIdx.x * BLOCK + threadIdx.x, [&](int h, int y, int q) {
        const bool y_fluid = (unsigned)map[y] < (unsigned)nf;
        const int x = h + L.off(q);
        int xi, xj, xk;
        decode(L, x, xi, xj, xk);
        const bool x_inside = xi >= 1 && xi <= L.nx && xj >= 1 && xj <= L.ny && xk >= 1 && xk <= L.nz;
        const bool from_h = map[y] < 0 || (!y_fluid && x_inside && (unsigned)map[x] < (unsigned)nf);
        for (int c = 0; c < ncol; ++c) {
            double* a = dense + (long)(19 * c + q) * ps;
            if (from_h) a[y] = a[h];
            else a[h] = a[y];
        }
    });
}

// message of the last failed create call (contexts keep their own)
inline thread_local std::string g_create_error;

// record an error on a context (or, for a NULL context, as the create error) and return its code
inline int vfail(std::string* dst, int code, const char* fmt, va_list ap) {
    char buf[512];
    vsnprintf(buf, sizeof buf, fmt, ap);
    *(dst ? dst : &g_create_error) = buf;
    return code;
}
template <class Ctx>
int fail(Ctx* c, int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    const int r = vfail(c ? &c->err : nullptr, code, fmt, ap);
    va_end(ap);
    return r;
}
inline int fail(std::nullptr_t, int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    const int r = vfail(nullptr, code, fmt, ap);
    va_end(ap);
    return r;
}

}  // namespace mflbm_dev

#define HIPCHK(c, call)                                                                                      \
    do {                                                                                                     \
        hipError_t e_ = (call);                                                                              \
        if (e_ != hipSuccess) return fail((c), MFLBM_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

#define LAUNCH(c, kern, grid, ...)                                                                           \
    do {                                                                                                     \
        if ((grid) > 0) {                                                                                    \
            hipLaunchKernelGGL(kern, dim3(grid), dim3(BLOCK), 0, (c)->stream, __VA_ARGS__);                  \
            hipError_t e_ = hipGetLastError();                                                               \
            if (e_ != hipSuccess) return fail((c), MFLBM_ERR_HIP, "launch %s: %s", #kern, hipGetErrorString(e_)); \
        }                                                                                                    \
    } while (0)
