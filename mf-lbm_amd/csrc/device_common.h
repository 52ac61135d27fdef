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
#include "mflbm.h"

namespace mflbm_dev {

constexpr int XO = 15, YO = 3, ZO = 3;
constexpr int BLOCK = 256;
constexpr int NXCD = 8;   // gfx950: 8 accelerator complex dies, work-groups dispatched round-robin

struct Layout {
    int nx, ny, nz;
    int px, pxy;   // row and plane pitch (elements)
    long total;    // elements per dense array
    __host__ __device__ int at(int i, int j, int k) const { return (i + XO) + (j + YO) * px + (k + ZO) * pxy; }
    __host__ __device__ int off(int q) const { return lat::EX[q] + lat::EY[q] * px + lat::EZ[q] * pxy; }
};

inline Layout make_layout(int nx, int ny, int nz) {
    Layout L{};
    L.nx = nx;
    L.ny = ny;
    L.nz = nz;
    L.px = ((nx + 20 + 15) / 16) * 16;
    L.pxy = L.px * (ny + 8);
    L.total = (long)L.pxy * (nz + 8);
    return L;
}

// Logical block index with each XCD taking one contiguous range of blocks: node lists are in x-fastest
// order, so a die then works on a z-slab and stencil neighbours stay in its own L2.  The grid is launched
// as NXCD * per blocks.
__device__ __forceinline__ int xcd_block(int per) { return (int)(blockIdx.x % NXCD) * per + (int)(blockIdx.x / NXCD); }

inline unsigned nblocks(long n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }
// grid of a list kernel that uses xcd_block(): per = blocks per die
inline int xcd_per(long n) { return (int)((nblocks(n) + NXCD - 1) / NXCD); }

// copy between the reference's halo-h layout (host) and the padded device layout (one array)
template <typename T>
void pack_host(const Layout& L, int h, const T* src, T* dst_padded) {
    const int sx = L.nx + 2 * h, sy = L.ny + 2 * h;
    for (int k = 1 - h; k <= L.nz + h; ++k)
        for (int j = 1 - h; j <= L.ny + h; ++j) {
            const T* s = src + (long)(j + h - 1) * sx + (long)(k + h - 1) * sx * sy;
            std::memcpy(dst_padded + L.at(1 - h, j, k), s, sizeof(T) * sx);
        }
}
template <typename T>
void unpack_host(const Layout& L, int h, const T* src_padded, T* dst) {
    const int sx = L.nx + 2 * h, sy = L.ny + 2 * h;
    for (int k = 1 - h; k <= L.nz + h; ++k)
        for (int j = 1 - h; j <= L.ny + h; ++j) {
            T* d = dst + (long)(j + h - 1) * sx + (long)(k + h - 1) * sx * sy;
            std::memcpy(d, src_padded + L.at(1 - h, j, k), sizeof(T) * sx);
        }
}

// Node order of the compact fluid slots and of the node lists: y-tiles of `ty` padded rows, then z, then y,
// then x.  Each die sweeps its share of the tiles along z; a stencil's neighbour rows then come from a few
// tile planes that stay in the die's L2, where full x-y planes of a large lattice would not.  ty >= ny + 8
// gives the plain x-fastest order.
struct TileOrder {
    int ty;
    Layout L;
    long key(int o) const {   // position of padded offset o in the order
        const int ip = o % L.px, jp = (o / L.px) % (L.ny + 8), kp = o / L.pxy;
        return ((long)(jp / ty) * (L.nz + 8) + kp) * ((long)ty * L.px) + (long)(jp % ty) * L.px + ip;
    }
    // visit the nodes of the box [lo, hi] (per axis, 1-based coordinates) in this order
    template <class F>
    void visit(int ilo, int ihi, int jlo, int jhi, int klo, int khi, F&& f) const {
        const int py = L.ny + 8;
        for (int jt = 0; jt * ty < py; ++jt)
            for (int k = klo; k <= khi; ++k)
                for (int jp = jt * ty; jp < py && jp < (jt + 1) * ty; ++jp) {
                    const int j = jp - YO;
                    if (j < jlo || j > jhi) continue;
                    for (int i = ilo; i <= ihi; ++i) f(i, j, k);
                }
    }
};

// rows per y-tile: MFLBM_TILE_Y in the environment, else 32 (the fastest of 4..32 and untiled on the 256^3
// two-phase and single-phase benchmarks, by 1.5 %)
inline int tile_rows() {
    const char* e = std::getenv("MFLBM_TILE_Y");
    const int v = e ? std::atoi(e) : 32;
    return v > 0 ? v : 32;
}

// Compact node set of a wall array (padded layout, walls != 0 is not fluid).
//   open_z: the z faces carry inlet/outlet conditions (planes k = 0, 1, 2, nz-1, nz, nz+1 are addressed)
//   kper / jper: periodic exchange along z / y (planes 0, 1, n, n+1)
//   extra_z: further z planes addressed by a boundary condition (e.g. a porous plate and its neighbours)
struct CompactSet {
    std::vector<int> cidx;   // padded offset -> compact index or -1
    std::vector<int> cpos;   // compact index -> padded offset
    int nfl = 0;             // interior fluid nodes (the first nfl entries)
};

inline void build_compact(const Layout& L, const TileOrder& T, const std::vector<int8_t>& wp, bool open_z, bool kper,
                          bool jper, const std::vector<int>& extra_z, CompactSet& cs) {
    cs.cidx.assign(L.total, -1);
    cs.cpos.clear();
    cs.cpos.reserve((size_t)L.nx * L.ny * L.nz);
    T.visit(1, L.nx, 1, L.ny, 1, L.nz, [&](int i, int j, int k) {
        const int o = L.at(i, j, k);
        if (wp[o] == 0) {
            cs.cidx[o] = (int)cs.cpos.size();
            cs.cpos.push_back(o);
        }
    });
    cs.nfl = (int)cs.cpos.size();
    std::vector<uint8_t> need(L.total, 0);
    for (int m = 0; m < cs.nfl; ++m)
        for (int q = 1; q < 19; ++q) need[cs.cpos[m] + L.off(q)] = 1;
    auto plane_z = [&](int k) {
        for (int j = 0; j <= L.ny + 1; ++j)
            for (int i = 0; i <= L.nx + 1; ++i) need[L.at(i, j, k)] = 1;
    };
    auto plane_y = [&](int j) {
        for (int k = 0; k <= L.nz + 1; ++k)
            for (int i = 0; i <= L.nx + 1; ++i) need[L.at(i, j, k)] = 1;
    };
    if (open_z)
        for (int k : {0, 1, 2, L.nz - 1, L.nz, L.nz + 1}) plane_z(k);
    if (kper)
        for (int k : {0, 1, L.nz, L.nz + 1}) plane_z(k);
    if (jper)
        for (int j : {0, 1, L.ny, L.ny + 1}) plane_y(j);
    for (int k : extra_z)
        if (k >= 0 && k <= L.nz + 1) plane_z(k);
    for (int k = 0; k <= L.nz + 1; ++k)
        for (int j = 0; j <= L.ny + 1; ++j)
            for (int i = 0; i <= L.nx + 1; ++i) {
                const int o = L.at(i, j, k);
                if (need[o] && cs.cidx[o] < 0) {
                    cs.cidx[o] = (int)cs.cpos.size();
                    cs.cpos.push_back(o);
                }
            }
}

// Link slots.  In the AA update the slot (x + e_r, opp(r)) of a non-fluid neighbour of fluid node x is read and
// written only by x: pulled as x's incoming population opp(r) at an odd step and refilled with x's outgoing
// population r.  Such slots are kept behind the compact slots of slice opp(r), packed per direction in
// fluid-node order, so the 64 lanes of a wave find their slots of one direction contiguous.  Per wave w of 64
// consecutive fluid nodes and direction r (entry 18 w + r - 1):
//   wmask: bit l set when fluid node 64 w + l has a non-fluid neighbour in direction r
//   wbase: number of such links of direction r among the fluid nodes before the wave
//   link slot of (64 w + l, r) = nc + wbase + popcount(wmask & (2^l - 1))
struct LinkSet {
    std::vector<unsigned long long> wmask;
    std::vector<int> wbase;
    long nlink[19] = {0};
};

inline void build_links(const Layout& L, const CompactSet& cs, LinkSet& ls) {
    const int nfl = cs.nfl;
    const int nwaves = (nfl + 63) / 64;
    ls.wmask.assign((size_t)nwaves * 18 + 1, 0ull);
    ls.wbase.assign((size_t)nwaves * 18 + 1, 0);
    for (int r = 0; r < 19; ++r) ls.nlink[r] = 0;
    for (int m = 0; m < nfl; ++m) {
        const size_t w = (size_t)(m >> 6);
        if ((m & 63) == 0)
            for (int r = 1; r < 19; ++r) ls.wbase[18 * w + r - 1] = (int)ls.nlink[r];
        for (int r = 1; r < 19; ++r)
            if (cs.cidx[cs.cpos[m] + L.off(r)] >= nfl) {
                ls.wmask[18 * w + r - 1] |= 1ull << (m & 63);
                ls.nlink[r]++;
            }
    }
}

// link slot of fluid node x (compact index) in direction r, or -1 when x + e_r is a fluid node
__device__ __forceinline__ int link_slot(int nc, const int* wbase, const unsigned long long* wmask, int x, int r) {
    const int e = 18 * (x >> 6) + r - 1;
    const unsigned long long wm = wmask[e];
    const unsigned l = (unsigned)x & 63u;
    if (!((wm >> l) & 1ull)) return -1;
    return nc + wbase[e] + __popcll(wm & ((1ull << l) - 1ull));
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
