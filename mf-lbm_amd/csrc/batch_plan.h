// Host-side plan of a batch of single-phase samples (mflbm_spb_commit): per sample the padded layout, the
// compact node set and link-slot tables, the slice offsets inside the one shared population buffer and the
// wave tables of the batched kernels.  Free of HIP, so CPU checks (AddressSanitizer builds) run it on real
// sample sets.
#pragma once

#include <algorithm>
#include <string>
#include <vector>

#include "layout.h"

namespace mflbm_dev {

struct BatchSampleIn {
    int nx, ny, nz, jper;
    const int8_t* walls_h2;   // (nx + 4)(ny + 4)(nz + 4), halo 2
};

struct BatchPlan {
    std::vector<Layout> L;
    std::vector<CompactSet> cs;
    std::vector<LinkSet> ls;
    std::vector<int> nf, nc;
    std::vector<int> wave_first, nwave;   // per sample: first global wave and wave count (64 fluid nodes each)
    std::vector<int> wave_sample;         // per global wave: its sample (at least one entry)
    std::vector<long> elem_off;           // per sample: first element of its slices in the shared buffer
    std::vector<long> soff;               // 20 per sample: slice q spans [soff[q], soff[q + 1]) inside the sample
    long elems = 0;                       // elements of the shared buffer
    long maxdense = 0;                    // largest padded lattice (staging of uploads and downloads)
    long nfluid = 0;
    int nwaves = 0;
};

// Lays the batch out (z periodic, y periodic where jper, one node order of `tile` rows per y tile); esz is the
// size of a stored population (4 or 8 bytes).  Returns an empty string, or why the batch cannot be laid out.
// every sample is periodic along z, and along y where jper; z is folded into the neighbour map (layout.h) where it
// is the only periodic axis
inline Fold batch_fold(int jper) { return Fold{false, jper != 1}; }

inline std::string plan_batch(const std::vector<BatchSampleIn>& in, int tile, long esz, BatchPlan& P) {
    const int n = (int)in.size();
    P = BatchPlan{};
    if (n == 0) return "no samples";
    P.L.resize(n);
    P.cs.resize(n);
    P.ls.resize(n);
    P.nf.resize(n);
    P.nc.resize(n);
    P.wave_first.resize(n);
    P.nwave.resize(n);
    P.elem_off.resize(n);
    P.soff.assign(20L * n, 0);
    for (int s = 0; s < n; ++s) {
        const BatchSampleIn& h = in[s];
        if (h.nx < 3 || h.ny < 3 || h.nz < 3 || !h.walls_h2) return "sample " + std::to_string(s) + ": bad shape or walls";
        const Layout L = make_layout(h.nx, h.ny, h.nz);
        if (L.total > (1L << 31) - 1) return "sample " + std::to_string(s) + " too large for 32-bit offsets";
        P.L[s] = L;
        std::vector<int8_t> wp(L.total, 1);
        pack_host(L, 2, h.walls_h2, wp.data());
        const TileOrder T{tile, L};
        build_compact(L, T, wp, false, true, h.jper == 1, {}, P.cs[s], batch_fold(h.jper));
        build_links(L, P.cs[s], P.ls[s]);
        P.nf[s] = P.cs[s].nfl;
        P.nc[s] = (int)P.cs[s].cpos.size();
        P.nfluid += P.cs[s].nfl;
        P.wave_first[s] = P.nwaves;
        P.nwave[s] = (P.cs[s].nfl + 63) / 64;
        P.nwaves += P.nwave[s];
        long* so = &P.soff[20L * s];
        for (int q = 0; q < 19; ++q) {
            const long len = (((long)P.nc[s] + (q ? P.ls[s].nlink[lat::OPP[q]] : 0) + 63) / 64) * 64;
            if (len * esz >= (1L << 32)) return "sample " + std::to_string(s) + ": slice exceeds 32-bit byte offsets";
            so[q + 1] = so[q] + len;
        }
        P.elem_off[s] = P.elems;
        P.elems += so[19];
        P.maxdense = std::max(P.maxdense, L.total);
    }
    P.wave_sample.assign((size_t)std::max(P.nwaves, 1), 0);
    for (int s = 0; s < n; ++s)
        for (int w = 0; w < P.nwave[s]; ++w) P.wave_sample[(size_t)P.wave_first[s] + w] = s;
    return "";
}

}  // namespace mflbm_dev
