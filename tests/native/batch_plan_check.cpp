// CPU check of the batched single-phase layout (mf-lbm_amd/csrc/batch_plan.h), built with AddressSanitizer and
// UndefinedBehaviorSanitizer by tests/test_batch_plan_asan.py.  Reads a sample set, plans it for every tile
// height the autotuner tries and both storage sizes, then replays on the host every index the batched kernels
// compute from the plan and checks it against the buffer it addresses:
//   * the lattice update: a wave's sample from the wave table, its fluid node, the slot of each direction in
//     the own node (even step) and in the neighbour or the node's link slot (odd step, AA pull/push);
//   * the folded periodic halo of the samples periodic along z only: every exchanged halo node (z = 0, nz + 1)
//     holds its periodic image's compact index, and the image lies in the grid; the periodic exchange planes
//     z = 0, 1, nz, nz + 1 of the other samples: every addressed node is stored;
//   * the monitor tables (flux per wave, per-sample reduction ranges).
// Input: int32 n, then per sample int32 nx, ny, nz, jper and (nx + 4)(ny + 4)(nz + 4) int8 walls with halo 2.
// Exit status 0 and "ok" on success; a message and status 1 on the first violation.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "batch_plan.h"

using namespace mflbm_dev;

static int bad(const char* what, int s, long a, long b) {
    std::printf("violation: %s (sample %d: %ld vs %ld)\n", what, s, a, b);
    return 1;
}

// the slot of direction q at padded offset o as the exchange kernels resolve it (support node -> owner's link)
static long slot_of(const BatchPlan& P, int s, int q, int o) {
    const Layout& L = P.L[s];
    const CompactSet& cs = P.cs[s];
    const LinkSet& ls = P.ls[s];
    const int c = cs.cidx.at(o);
    if (c >= cs.nfl) {
        const int x = cs.cidx.at(o + L.off(q));
        if (x >= 0 && x < cs.nfl) {
            const int r = lat::OPP[q];
            const size_t e = 18 * (size_t)(x >> 6) + r - 1;
            const unsigned long long wm = ls.wmask.at(e);
            const unsigned l = (unsigned)x & 63u;
            if ((wm >> l) & 1ull)
                return (long)P.nc[s] + ls.wbase.at(e) + __builtin_popcountll(wm & ((1ull << l) - 1ull));
        }
    }
    return c;
}

static int check(const BatchPlan& P, long esz, const std::vector<BatchSampleIn>& in) {
    const int n = (int)P.L.size();
    if ((int)P.wave_sample.size() < std::max(P.nwaves, 1)) return bad("wave table size", -1, (long)P.wave_sample.size(), P.nwaves);
    long elems = 0;
    for (int s = 0; s < n; ++s) {
        const Layout& L = P.L[s];
        const CompactSet& cs = P.cs[s];
        const LinkSet& ls = P.ls[s];
        if (P.elem_off.at(s) != elems) return bad("element offset", s, P.elem_off.at(s), elems);
        const long* so = &P.soff.at(20L * s);
        for (int q = 0; q < 19; ++q) {
            const long len = so[q + 1] - so[q];
            if (len < P.nc[s]) return bad("slice shorter than the compact slots", s, len, P.nc[s]);
            if (len * esz >= (1L << 32)) return bad("slice beyond 32-bit byte offsets", s, len, q);
        }
        elems += so[19];
        if ((long)cs.cidx.size() != L.total) return bad("map size", s, (long)cs.cidx.size(), L.total);
        // wave table: the waves of sample s cover its fluid nodes
        for (int w = 0; w < P.nwave.at(s); ++w) {
            const int gw = P.wave_first.at(s) + w;
            if (P.wave_sample.at(gw) != s) return bad("wave sample", s, gw, P.wave_sample.at(gw));
        }
        // lattice update, both parities, every fluid node
        for (int m = 0; m < cs.nfl; ++m) {
            const int o = cs.cpos.at(m);
            if (o < 0 || o >= L.total) return bad("fluid position", s, o, L.total);
            for (int r = 1; r < 19; ++r) {
                const size_t e = 18 * (size_t)(m >> 6) + r - 1;
                const unsigned long long wm = ls.wmask.at(e);
                const unsigned lane = (unsigned)m & 63u;
                long idx;
                const int q = lat::OPP[r];   // the slot (x + e_r, opp(r)) lives in slice opp(r)
                if ((wm >> lane) & 1ull) {
                    idx = (long)P.nc[s] + ls.wbase.at(e) + __builtin_popcountll(wm & ((1ull << lane) - 1ull));
                } else {
                    const long on = o + L.off(r);
                    if (on < 0 || on >= L.total) return bad("neighbour offset", s, on, L.total);
                    idx = cs.cidx.at(on);
                    if (idx < 0 || idx >= cs.nfl) return bad("neighbour without a link slot is not fluid", s, m, idx);
                }
                if (idx < 0 || idx >= so[q + 1] - so[q]) return bad("odd-step slot outside its slice", s, idx, so[q + 1] - so[q]);
            }
        }
        // periodic exchange planes of the samples that keep their exchange (periodic along y and z)
        auto plane_ok = [&](int i, int j, int k) -> bool {
            const int o = L.at(i, j, k);
            if (o < 0 || o >= L.total) return false;
            for (int q = 1; q < 19; ++q) {
                const long id = slot_of(P, s, q, o);
                if (id < 0 || id >= so[q + 1] - so[q]) return false;
            }
            return true;
        };
        if (!cs.fold.z)
            for (int k : {0, 1, L.nz, L.nz + 1})
                for (int j = 1; j <= L.ny; ++j)
                    for (int i = 1; i <= L.nx; ++i)
                        if (!plane_ok(i, j, k)) return bad("z exchange plane node not stored", s, i + 1000L * j, k);
        // folded periodic halo: each exchanged halo node reads its image's compact index
        if (cs.fold.y || cs.fold.z != (in.at(s).jper != 1)) return bad("fold of the periodic axes", s, cs.fold.z, cs.fold.y);
        for (int k = 0; k <= L.nz + 1; ++k)
            for (int j = 0; j <= L.ny + 1; ++j)
                for (int i = 1; i <= L.nx; ++i) {
                    int io;
                    if (!fold_image(L, cs.fold, i, j, k, io)) continue;
                    if (io < 0 || io >= L.total) return bad("fold image outside the grid", s, io, L.total);
                    if (cs.cidx.at(L.at(i, j, k)) != cs.cidx.at(io))
                        return bad("folded halo node without its image's index", s, i + 1000L * j, k);
                }
        if (ls.wmask.size() < 18 * (size_t)((cs.nfl + 63) / 64)) return bad("link mask table", s, (long)ls.wmask.size(), cs.nfl);
    }
    if (elems != P.elems) return bad("total elements", -1, elems, P.elems);
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::printf("usage: %s samples.bin\n", argv[0]);
        return 2;
    }
    FILE* fh = std::fopen(argv[1], "rb");
    if (!fh) return 2;
    int n = 0;
    if (std::fread(&n, 4, 1, fh) != 1 || n < 1) return 2;
    std::vector<std::vector<int8_t>> walls(n);
    std::vector<BatchSampleIn> in(n);
    for (int s = 0; s < n; ++s) {
        int hdr[4];
        if (std::fread(hdr, 4, 4, fh) != 4) return 2;
        walls[s].resize((size_t)(hdr[0] + 4) * (hdr[1] + 4) * (hdr[2] + 4));
        if (std::fread(walls[s].data(), 1, walls[s].size(), fh) != walls[s].size()) return 2;
        in[s] = BatchSampleIn{hdr[0], hdr[1], hdr[2], hdr[3], walls[s].data()};
    }
    std::fclose(fh);
    long total_nodes = 0;
    for (int tile : {8, 16, 32, 64})
        for (long esz : {4L, 8L}) {
            BatchPlan P;
            const std::string why = plan_batch(in, tile, esz, P);
            if (!why.empty()) {
                std::printf("plan refused: %s\n", why.c_str());
                return 1;
            }
            if (check(P, esz, in)) {
                std::printf("(tile %d, %ld-byte storage)\n", tile, esz);
                return 1;
            }
            total_nodes = P.nfluid;
        }
    std::printf("ok: %d samples, %ld fluid nodes\n", n, total_nodes);
    return 0;
}
