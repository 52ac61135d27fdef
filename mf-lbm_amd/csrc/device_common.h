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
// as NXCD * per blocks.
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
