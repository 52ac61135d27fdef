// Host-side layout of the gfx950 paths, free of any HIP dependency so that it also builds for CPU-only checks:
// the padded dense layout, the compact node set the populations are stored on and the per-wave link-slot
// tables.  See device_common.h for the description of the layouts.
#pragma once

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "lattice.h"

namespace mflbm_dev {

constexpr int XO = 15, YO = 3, ZO = 3;

struct Layout {
    int nx, ny, nz;
    int px, pxy;   // row and plane pitch (elements)
    long total;    // elements per dense array
    MFLBM_HD int at(int i, int j, int k) const { return (i + XO) + (j + YO) * px + (k + ZO) * pxy; }
    MFLBM_HD int off(int q) const { return lat::EX[q] + lat::EY[q] * px + lat::EZ[q] * pxy; }
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

// A periodic axis folded into the neighbour map.  On one domain periodic along one axis only, the reference's
// self-exchange (Mpi.F90:73-448) keeps every exchanged halo slot equal to the slot of its periodic image: after an
// even step the pull copies image -> halo, after an odd step the push copies halo -> image, between the two no
// node touches both, and each image slot has one halo slot.  A folded axis therefore maps the exchanged halo
// nodes (x interior, one layer) to the compact index of their image, and the lattice update pulls from and pushes
// to the image directly: the exchange launches disappear.  With both y and z periodic, a slot of an edge node has
// three halo aliases (face, face, edge) that the face and edge exchanges leave unequal after an odd step, so
// such domains keep their exchange.
struct Fold {
    bool y = false, z = false;
};
MFLBM_HD inline void decode(const Layout& L, int o, int& i, int& j, int& k) {
    const int ip = o % L.px, jp = (o / L.px) % (L.ny + 8), kp = o / L.pxy;
    i = ip - XO;
    j = jp - YO;
    k = kp - ZO;
}
// the periodic image of (i, j, k) when that node is an exchanged halo node of a folded axis
MFLBM_HD inline bool fold_image(const Layout& L, const Fold& F, int i, int j, int k, int& io) {
    if (i < 1 || i > L.nx) return false;
    const bool gy = j == 0 || j == L.ny + 1, gz = k == 0 || k == L.nz + 1;
    if (j < 0 || j > L.ny + 1 || k < 0 || k > L.nz + 1) return false;
    if (gy && gz) {
        if (!(F.y && F.z)) return false;
    } else if (gz) {
        if (!F.z) return false;
    } else if (gy) {
        if (!F.y) return false;
    } else {
        return false;
    }
    const int jj = gy ? (j == 0 ? L.ny : 1) : j, kk = gz ? (k == 0 ? L.nz : 1) : k;
    io = L.at(i, jj, kk);
    return true;
}

// Compact node set of a wall array (padded layout, walls != 0 is not fluid).
//   open_z: the z faces carry inlet/outlet conditions (planes k = 0, 1, 2, nz-1, nz, nz+1 are addressed)
//   kper / jper: periodic exchange along z / y (planes 0, 1, n, n+1), unless the axis is folded
//   extra_z: further z planes addressed by a boundary condition (e.g. a porous plate and its neighbours)
//   fold: periodic axes folded into the map (Fold); a folded halo node takes its image's compact index
struct CompactSet {
    std::vector<int> cidx;   // padded offset -> compact index or -1
    std::vector<int> cpos;   // compact index -> padded offset
    int nfl = 0;             // interior fluid nodes (the first nfl entries)
    Fold fold;
};

inline void build_compact(const Layout& L, const TileOrder& T, const std::vector<int8_t>& wp, bool open_z, bool kper,
                          bool jper, const std::vector<int>& extra_z, CompactSet& cs, Fold fold = Fold{}) {
    cs.cidx.assign(L.total, -1);
    cs.cpos.clear();
    cs.cpos.reserve((size_t)L.nx * L.ny * L.nz);
    cs.fold = fold;
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
        for (int q = 1; q < 19; ++q) {
            int o = cs.cpos[m] + L.off(q), i, j, k, io;
            decode(L, o, i, j, k);
            if (fold_image(L, fold, i, j, k, io)) o = io;
            need[o] = 1;
        }
    if (fold.z) kper = false;
    if (fold.y) jper = false;
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
    if (fold.y || fold.z)
        for (int k = 0; k <= L.nz + 1; ++k)
            for (int j = 0; j <= L.ny + 1; ++j)
                for (int i = 1; i <= L.nx; ++i) {
                    int io;
                    if (fold_image(L, fold, i, j, k, io)) cs.cidx[L.at(i, j, k)] = cs.cidx[io];
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

}  // namespace mflbm_dev
