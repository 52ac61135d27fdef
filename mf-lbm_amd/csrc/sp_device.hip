// gfx950 implementation of MF-LBM's single-phase time step (D3Q19 MRT, singlephase_3D/0.src) and its
// C ABI (the mflbm_sp_* entry points of mflbm.h).
//
// precision 0 keeps FP64 populations (the reference arithmetic), precision 1 FP32 deviations from the rest
// equilibrium (sp_kernels.h); the open boundaries and macroscopic fields below run in FP64 in both modes.
//
// Storage follows the two-phase path (device_common.h): the 19 population arrays are compact over the
// interior fluid nodes plus the support nodes the update, the open boundaries and the periodic exchange
// address; macroscopic fields and walls live on the padded dense grid.  One thread updates one fluid node;
// the AA pattern makes every node read and write the same slot set, so the update has no races and the
// result does not depend on the thread order (Kernel.F90:5-400).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "collide_sp.h"
#include "device_common.h"
#include "lattice.h"
#include "mflbm.h"
#include "sp_kernels.h"

using namespace mflbm_dev;

namespace {

using namespace mflbm_sp;

// ------------------------------------------------------------------------------------------------------
// open boundaries (Boundary.F90), one thread per (i, j) of the face; `x * (1 - w) + y * w` blends are kept
// literally (a solid face node keeps its value)
// ------------------------------------------------------------------------------------------------------

// inlet_bounce_back_velocity_BC_before_odd (Boundary.F90:5-37)
template <class T>
__global__ void k_sp_inlet_velocity_before(SDev<T> d, double relaxation) {
    const Layout L = d.L;
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= L.nx * L.ny) return;
    const int i = t % L.nx + 1, j = t / L.nx + 1;
    const int wi = d.walls[L.at(i, j, 1)];
    const double om = (double)(1 - wi), wd = (double)wi;
    const double s = d.w_in[plane_at(d, i, j)] * relaxation;
    const double c1 = 6.0 * lat::W1, c2 = 6.0 * lat::W2;
    const SPop<T> A = spop(d);
    int o;
    o = L.at(i, j, 0);     A(5, o) = (A(6, L.at(i, j, 1)) + c1 * s) * om + A(5, o) * wd;
    o = L.at(i - 1, j, 0); A(11, o) = (A(14, L.at(i, j, 1)) + c2 * s) * om + A(11, o) * wd;
    o = L.at(i + 1, j, 0); A(12, o) = (A(13, L.at(i, j, 1)) + c2 * s) * om + A(12, o) * wd;
    o = L.at(i, j - 1, 0); A(15, o) = (A(18, L.at(i, j, 1)) + c2 * s) * om + A(15, o) * wd;
    o = L.at(i, j + 1, 0); A(16, o) = (A(17, L.at(i, j, 1)) + c2 * s) * om + A(16, o) * wd;
}
// inlet_bounce_back_velocity_BC_after_odd (Boundary.F90:40-70)
template <class T>
__global__ void k_sp_inlet_velocity_after(SDev<T> d, double relaxation) {
    const Layout L = d.L;
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= L.nx * L.ny) return;
    const int i = t % L.nx + 1, j = t / L.nx + 1;
    const int wi = d.walls[L.at(i, j, 1)];
    const double om = (double)(1 - wi), wd = (double)wi;
    const double s = d.w_in[plane_at(d, i, j)] * relaxation;
    const double c1 = 6.0 * lat::W1, c2 = 6.0 * lat::W2;
    const SPop<T> A = spop(d);
    const int o = L.at(i, j, 1);
    A(6, o) = (A(5, L.at(i, j, 0)) + c1 * s) * om + A(6, o) * wd;
    A(13, o) = (A(12, L.at(i + 1, j, 0)) + c2 * s) * om + A(13, o) * wd;
    A(14, o) = (A(11, L.at(i - 1, j, 0)) + c2 * s) * om + A(14, o) * wd;
    A(17, o) = (A(16, L.at(i, j + 1, 0)) + c2 * s) * om + A(17, o) * wd;
    A(18, o) = (A(15, L.at(i, j - 1, 0)) + c2 * s) * om + A(18, o) * wd;
}
// inlet_Zou_He_pressure_BC_before_odd (Boundary.F90:76-132)
template <class T>
__global__ void k_sp_inlet_pressure_before(SDev<T> d, double rho_in, double relaxation) {
    const Layout L = d.L;
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= L.nx * L.ny) return;
    const int i = t % L.nx + 1, j = t / L.nx + 1;
    const int wi = d.walls[L.at(i, j, 1)];
    const double om = (double)(1 - wi), wd = (double)wi;
    const SPop<T> A = spop(d);
    const double s = (rho_in - (A(0, L.at(i, j, 1)) + A(1, L.at(i - 1, j, 1)) + A(2, L.at(i + 1, j, 1)) +
                                A(3, L.at(i, j - 1, 1)) + A(4, L.at(i, j + 1, 1)) + A(7, L.at(i - 1, j - 1, 1)) +
                                A(8, L.at(i + 1, j - 1, 1)) + A(9, L.at(i - 1, j + 1, 1)) + A(10, L.at(i + 1, j + 1, 1)) +
                                2.0 * (A(6, L.at(i, j, 2)) + A(14, L.at(i + 1, j, 2)) + A(13, L.at(i - 1, j, 2)) +
                                       A(18, L.at(i, j + 1, 2)) + A(17, L.at(i, j - 1, 2))))) * relaxation;
    const double tnx = 0.5 * (A(1, L.at(i - 1, j, 1)) + A(7, L.at(i - 1, j - 1, 1)) + A(9, L.at(i - 1, j + 1, 1)) -
                              (A(2, L.at(i + 1, j, 1)) + A(8, L.at(i + 1, j - 1, 1)) + A(10, L.at(i + 1, j + 1, 1))));
    const double tny = 0.5 * (A(3, L.at(i, j - 1, 1)) + A(7, L.at(i - 1, j - 1, 1)) + A(8, L.at(i + 1, j - 1, 1)) -
                              (A(4, L.at(i, j + 1, 1)) + A(10, L.at(i + 1, j + 1, 1)) + A(9, L.at(i - 1, j + 1, 1))));
    int o;
    o = L.at(i, j, 0);     A(5, o) = (A(6, L.at(i, j, 2)) + lat::ONE3 * s) * om + A(5, o) * wd;
    o = L.at(i - 1, j, 0); A(11, o) = (A(14, L.at(i + 1, j, 2)) + lat::ONE6 * s - tnx) * om + A(11, o) * wd;
    o = L.at(i + 1, j, 0); A(12, o) = (A(13, L.at(i - 1, j, 2)) + lat::ONE6 * s + tnx) * om + A(12, o) * wd;
    o = L.at(i, j - 1, 0); A(15, o) = (A(18, L.at(i, j + 1, 2)) + lat::ONE6 * s - tny) * om + A(15, o) * wd;
    o = L.at(i, j + 1, 0); A(16, o) = (A(17, L.at(i, j - 1, 2)) + lat::ONE6 * s + tny) * om + A(16, o) * wd;
}
// inlet_Zou_He_pressure_BC_after_odd (Boundary.F90:135-183)
template <class T>
__global__ void k_sp_inlet_pressure_after(SDev<T> d, double rho_in, double relaxation) {
    const Layout L = d.L;
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= L.nx * L.ny) return;
    const int i = t % L.nx + 1, j = t / L.nx + 1;
    const int wi = d.walls[L.at(i, j, 1)];
    const double om = (double)(1 - wi), wd = (double)wi;
    const SPop<T> A = spop(d);
    const int o = L.at(i, j, 1);
    const double s = (rho_in - (A(0, o) + A(2, o) + A(1, o) + A(4, o) + A(3, o) + A(8, o) + A(7, o) + A(10, o) + A(9, o) +
                                2.0 * (A(5, o) + A(11, o) + A(12, o) + A(15, o) + A(16, o)))) * relaxation;
    const double tnx = 0.5 * (A(2, o) + A(8, o) + A(10, o) - (A(1, o) + A(7, o) + A(9, o)));
    const double tny = 0.5 * (A(4, o) + A(9, o) + A(10, o) - (A(3, o) + A(8, o) + A(7, o)));
    A(6, o) = (A(5, o) + lat::ONE3 * s) * om + A(6, o) * wd;
    A(13, o) = (A(12, o) + lat::ONE6 * s + tnx) * om + A(13, o) * wd;
    A(14, o) = (A(11, o) + lat::ONE6 * s - tnx) * om + A(14, o) * wd;
    A(17, o) = (A(16, o) + lat::ONE6 * s + tny) * om + A(17, o) * wd;
    A(18, o) = (A(15, o) + lat::ONE6 * s - tny) * om + A(18, o) * wd;
}
// outlet_convective_BC_before_odd (Boundary.F90:192-230)
template <class T>
__global__ void k_sp_outlet_convective_before(SDev<T> d, double u_convec) {
    const Layout L = d.L;
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= L.nx * L.ny) return;
    const int i = t % L.nx + 1, j = t / L.nx + 1, nz = L.nz;
    const double temp = 1.0 / (1.0 + u_convec);
    const int wi = d.walls[L.at(i, j, nz)];
    const double om = (double)(1 - wi), wd = (double)wi;
    const int pc = plane_at(d, i, j);
    const long pl = (long)(L.nx + 2) * (L.ny + 2);
    // (slot, x offset, y offset) of the five unknown populations at the outlet halo
    const int QS[5] = {6, 13, 14, 17, 18};
    const int DI[5] = {0, -1, 1, 0, 0};
    const int DJ[5] = {0, 0, 0, -1, 1};
    const SPop<T> A = spop(d);
    for (int m = 0; m < 5; ++m) {
        const int o = L.at(i + DI[m], j + DJ[m], nz + 1);
        const int in = L.at(i + DI[m], j + DJ[m], nz);
        A(QS[m], o) = ((d.cf[pc + QS[m] * pl] + u_convec * A(QS[m], in)) * temp) * om + A(QS[m], o) * wd;
    }
    for (int m = 0; m < 5; ++m) d.cf[pc + QS[m] * pl] = A(QS[m], L.at(i + DI[m], j + DJ[m], nz + 1));
}
// outlet_convective_BC_after_odd (Boundary.F90:233-267)
template <class T>
__global__ void k_sp_outlet_convective_after(SDev<T> d, double u_convec) {
    const Layout L = d.L;
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= L.nx * L.ny) return;
    const int i = t % L.nx + 1, j = t / L.nx + 1, nz = L.nz;
    const double temp = 1.0 / (1.0 + u_convec);
    const int wi = d.walls[L.at(i, j, nz)];
    const double om = (double)(1 - wi), wd = (double)wi;
    const int pc = plane_at(d, i, j);
    const long pl = (long)(L.nx + 2) * (L.ny + 2);
    // slot written at z = nz and the history slot it relaxes toward
    const int QW[5] = {5, 11, 12, 15, 16};
    const int QH[5] = {6, 14, 13, 18, 17};
    const int o = L.at(i, j, nz), om1 = L.at(i, j, nz - 1);
    const SPop<T> A = spop(d);
    for (int m = 0; m < 5; ++m)
        A(QW[m], o) = ((d.cf[pc + QH[m] * pl] + u_convec * A(QW[m], om1)) * temp) * om + A(QW[m], o) * wd;
    for (int m = 0; m < 5; ++m) d.cf[pc + QH[m] * pl] = A(QW[m], o);
}
// outlet_Zou_He_pressure_BC_before_odd (Boundary.F90:276-330)
template <class T>
__global__ void k_sp_outlet_pressure_before(SDev<T> d, double rho_out) {
    const Layout L = d.L;
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= L.nx * L.ny) return;
    const int i = t % L.nx + 1, j = t / L.nx + 1, nz = L.nz;
    const int wi = d.walls[L.at(i, j, nz)];
    const double om = 1.0 - (double)wi, wd = (double)wi;
    const SPop<T> A = spop(d);
    const int a0 = L.at(i, j, nz), am = L.at(i - 1, j, nz), ap = L.at(i + 1, j, nz), bm = L.at(i, j - 1, nz),
              bp = L.at(i, j + 1, nz), mm = L.at(i - 1, j - 1, nz), pm = L.at(i + 1, j - 1, nz),
              mp = L.at(i - 1, j + 1, nz), pp = L.at(i + 1, j + 1, nz);
    const int c1 = L.at(i, j, nz - 1), c2 = L.at(i - 1, j, nz - 1), c3 = L.at(i + 1, j, nz - 1), c4 = L.at(i, j - 1, nz - 1),
              c5 = L.at(i, j + 1, nz - 1);
    const double s = (A(0, a0) + A(1, am) + A(2, ap) + A(3, bm) + A(4, bp) + A(7, mm) + A(8, pm) + A(9, mp) + A(10, pp) +
                      2.0 * (A(5, c1) + A(11, c2) + A(12, c3) + A(15, c4) + A(16, c5))) - rho_out;
    const double tnx = 0.5 * (A(1, am) + A(7, mm) + A(9, mp) - (A(2, ap) + A(8, pm) + A(10, pp)));
    const double tny = 0.5 * (A(3, bm) + A(7, mm) + A(8, pm) - (A(4, bp) + A(10, pp) + A(9, mp)));
    int o;
    o = L.at(i, j, nz + 1);     A(6, o) = (A(5, c1) - lat::ONE3 * s) * om + A(6, o) * wd;
    o = L.at(i - 1, j, nz + 1); A(13, o) = (A(12, c3) - lat::ONE6 * s - tnx) * om + A(13, o) * wd;
    o = L.at(i + 1, j, nz + 1); A(14, o) = (A(11, c2) - lat::ONE6 * s + tnx) * om + A(14, o) * wd;
    o = L.at(i, j - 1, nz + 1); A(17, o) = (A(16, c5) - lat::ONE6 * s - tny) * om + A(17, o) * wd;
    o = L.at(i, j + 1, nz + 1); A(18, o) = (A(15, c4) - lat::ONE6 * s + tny) * om + A(18, o) * wd;
}
// outlet_Zou_He_pressure_BC_after_odd (Boundary.F90:333-389)
template <class T>
__global__ void k_sp_outlet_pressure_after(SDev<T> d, double rho_out) {
    const Layout L = d.L;
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= L.nx * L.ny) return;
    const int i = t % L.nx + 1, j = t / L.nx + 1, nz = L.nz;
    const int wi = d.walls[L.at(i, j, nz)];
    const double om = 1.0 - (double)wi, wd = (double)wi;
    const SPop<T> A = spop(d);
    const int o = L.at(i, j, nz);
    const double s = (A(0, o) + A(2, o) + A(1, o) + A(4, o) + A(3, o) + A(8, o) + A(7, o) + A(10, o) + A(9, o) +
                      2.0 * (A(6, o) + A(14, o) + A(13, o) + A(18, o) + A(17, o))) - rho_out;
    const double tnx = 0.5 * (A(2, o) + A(8, o) + A(10, o) - (A(1, o) + A(7, o) + A(9, o)));
    const double tny = 0.5 * (A(4, o) + A(10, o) + A(9, o) - (A(3, o) + A(7, o) + A(8, o)));
    A(5, o) = (A(6, o) - lat::ONE3 * s) * om + A(5, o) * wd;
    A(11, o) = (A(14, o) - lat::ONE6 * s + tnx) * om + A(11, o) * wd;
    A(12, o) = (A(13, o) - lat::ONE6 * s - tnx) * om + A(12, o) * wd;
    A(15, o) = (A(18, o) - lat::ONE6 * s + tny) * om + A(15, o) * wd;
    A(16, o) = (A(17, o) - lat::ONE6 * s - tny) * om + A(16, o) * wd;
}

// ------------------------------------------------------------------------------------------------------
// macroscopic fields and monitor reductions
// ------------------------------------------------------------------------------------------------------

// monitor (Monitor.F90:20-48), one block per z layer: fl(k) = sum of w, pre(k) = sum of rho over the layer,
// and the layer's max of u^2+v^2+w^2.  The sums are tree reductions (the reference adds in x-fastest order).
__global__ void __launch_bounds__(BLOCK) k_sp_monitor_layers(Layout L, const double* u, const double* v, const double* w,
                                                             const double* rho, double* out) {
    const int k = blockIdx.x + 1;
    __shared__ double s_fl[BLOCK], s_pre[BLOCK], s_um[BLOCK];
    double fl = 0.0, pre = 0.0, um = 0.0;
    for (int t = threadIdx.x; t < L.nx * L.ny; t += BLOCK) {
        const int o = L.at(t % L.nx + 1, t / L.nx + 1, k);
        fl += w[o];
        pre += rho[o];
        const double q = u[o] * u[o] + v[o] * v[o] + w[o] * w[o];
        if (um < q) um = q;
    }
    s_fl[threadIdx.x] = fl;
    s_pre[threadIdx.x] = pre;
    s_um[threadIdx.x] = um;
    __syncthreads();
    for (int h = BLOCK / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            s_fl[threadIdx.x] += s_fl[threadIdx.x + h];
            s_pre[threadIdx.x] += s_pre[threadIdx.x + h];
            if (s_um[threadIdx.x] < s_um[threadIdx.x + h]) s_um[threadIdx.x] = s_um[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[k - 1] = s_fl[0];
        out[L.nz + k - 1] = s_pre[0];
        out[2 * L.nz + k - 1] = s_um[0];
    }
}

bool sp_open_z(const mflbm_sp_params& p) { return p.kper == 0 && p.wall_z_min == 0 && p.wall_z_max == 0; }
// the periodic axis folded into the neighbour map (layout.h): the one periodic axis of a domain periodic along z
// only, or along y only without an inlet/outlet face (which addresses the y halo rows)
Fold sp_fold_of(const mflbm_sp_params& p) {
    return Fold{p.jper == 1 && p.kper == 0 && !sp_open_z(p), p.kper == 1 && p.jper == 0};
}

}  // namespace

struct mflbm_sp_ctx {
    mflbm_sp_params p{};
    SPCollideParams cp{};
    SPCollideParamsT<float> cp32{};
    bool fp32 = false;           // precision 1 or 2: FP32 deviation storage
    bool mixed = false;          // precision 2: the FP32 deviations collided in FP64
    bool stepped = false;        // a step ran since the state was uploaded or built
    Layout L{};
    SDev<double> d{};            // slice pointers of the FP64 mode; the other fields are shared by both views
    SDev<float> d32{};
    hipStream_t stream = nullptr;
    long ps = 0;                 // stride of the dense arrays
    double* dense_pdf = nullptr; // 19 dense arrays: upload/download staging
    double* fields = nullptr;    // u, v, w, rho
    int8_t* walls = nullptr;
    double* planes = nullptr;    // w_in, cf (19)
    double* red = nullptr;       // monitor layer results (3*nz)
    void* compact = nullptr;     // 19 population slices (compact slots + link slots), double or float
    int* map = nullptr;
    int* cpos = nullptr;
    unsigned long long* wmask = nullptr;
    int* wbase = nullptr;
    bool have_geom = false, have_state = false;
    bool timing = false;
    std::vector<hipEvent_t> ev;  // 3 events per timed step: start, after the lattice update, end
    size_t ev_used = 0;
    double t_sum[3] = {0, 0, 0};
    long t_cnt[3] = {0, 0, 0};
    std::string err;
};

namespace {

template <class T>
const SDev<T>& view(const mflbm_sp_ctx* c);
template <>
const SDev<double>& view<double>(const mflbm_sp_ctx* c) { return c->d; }
template <>
const SDev<float>& view<float>(const mflbm_sp_ctx* c) { return c->d32; }
template <class T>
const SPCollideParamsT<T>& rates(const mflbm_sp_ctx* c);
template <>
const SPCollideParamsT<double>& rates<double>(const mflbm_sp_ctx* c) { return c->cp; }
template <>
const SPCollideParamsT<float>& rates<float>(const mflbm_sp_ctx* c) { return c->cp32; }

template <class T>
int sp_launch_step_t(mflbm_sp_ctx* c, int ntime, hipEvent_t* ev) {
    const mflbm_sp_params& p = c->p;
    const Layout& L = c->L;
    const SDev<T>& d = view<T>(c);
    const bool odd = (ntime % 2) != 0;
    const int per = xcd_per(d.nf);
    if (c->mixed) {
        if (odd) LAUNCH(c, (k_sp_collide<true, T, double>), NXCD * per, d, c->cp, per);
        else LAUNCH(c, (k_sp_collide<false, T, double>), NXCD * per, d, c->cp, per);
    } else {
        if (odd) LAUNCH(c, (k_sp_collide<true, T, T>), NXCD * per, d, rates<T>(c), per);
        else LAUNCH(c, (k_sp_collide<false, T, T>), NXCD * per, d, rates<T>(c), per);
    }
    if (ev) HIPCHK(c, hipEventRecord(ev[1], c->stream));
    if (p.kper == 1 || p.jper == 1) {
        // exchange of the periodic axes not folded into the neighbour map
        const Fold F = sp_fold_of(p);
        const bool wz = p.kper == 1 && !F.z, wy = p.jper == 1 && !F.y, we = p.kper == 1 && p.jper == 1 && !(F.y && F.z);
        if (!odd) {
            if (wz) LAUNCH(c, k_sp_pull_z<T>, nblocks((long)L.nx * L.ny), d);
            if (wy) LAUNCH(c, k_sp_pull_y<T>, nblocks((long)L.nx * L.nz), d);
            if (we) LAUNCH(c, k_sp_pull_edge<T>, nblocks(L.nx), d);
        } else {
            if (wz) LAUNCH(c, k_sp_push_z<T>, nblocks((long)L.nx * L.ny), d);
            if (wy) LAUNCH(c, k_sp_push_y<T>, nblocks((long)L.nx * L.nz), d);
            if (we) LAUNCH(c, k_sp_push_edge<T>, nblocks(L.nx), d);
        }
    }
    if (sp_open_z(p)) {
        const unsigned gface = nblocks((long)L.nx * L.ny);
        if (!odd) {
            if (p.inlet_bc == 1) LAUNCH(c, k_sp_inlet_velocity_before<T>, gface, d, p.relaxation);
            else if (p.inlet_bc == 2) LAUNCH(c, k_sp_inlet_pressure_before<T>, gface, d, p.rho_in, p.relaxation);
            if (p.outlet_bc == 1) LAUNCH(c, k_sp_outlet_convective_before<T>, gface, d, p.uin_avg);
            else if (p.outlet_bc == 2) LAUNCH(c, k_sp_outlet_pressure_before<T>, gface, d, p.rho_out);
        } else {
            if (p.inlet_bc == 1) LAUNCH(c, k_sp_inlet_velocity_after<T>, gface, d, p.relaxation);
            else if (p.inlet_bc == 2) LAUNCH(c, k_sp_inlet_pressure_after<T>, gface, d, p.rho_in, p.relaxation);
            if (p.outlet_bc == 1) LAUNCH(c, k_sp_outlet_convective_after<T>, gface, d, p.uin_avg);
            else if (p.outlet_bc == 2) LAUNCH(c, k_sp_outlet_pressure_after<T>, gface, d, p.rho_out);
        }
    }
    return MFLBM_OK;
}

int sp_launch_step(mflbm_sp_ctx* c, int ntime) {
    c->stepped = true;
    hipEvent_t* ev = nullptr;
    if (c->timing) {
        if (c->ev_used + 3 > c->ev.size()) {
            for (int m = 0; m < 3; ++m) {
                hipEvent_t e;
                HIPCHK(c, hipEventCreate(&e));
                c->ev.push_back(e);
            }
        }
        ev = &c->ev[c->ev_used];
        c->ev_used += 3;
        HIPCHK(c, hipEventRecord(ev[0], c->stream));
    }
    const int rc = c->fp32 ? sp_launch_step_t<float>(c, ntime, ev) : sp_launch_step_t<double>(c, ntime, ev);
    if (rc) return rc;
    if (ev) HIPCHK(c, hipEventRecord(ev[2], c->stream));
    return MFLBM_OK;
}

template <class T>
int sp_gather_t(mflbm_sp_ctx* c) {
    const SDev<T>& d = view<T>(c);
    LAUNCH(c, k_sp_gather<T>, nblocks(d.nc), d, (const double*)c->dense_pdf, c->ps);
    LAUNCH(c, (k_sp_links<false, T>), nblocks(d.nf), d, c->dense_pdf, c->ps);
    const Fold F = sp_fold_of(c->p);
    if (F.y || F.z) LAUNCH(c, k_sp_fold_upload<T>, fold_blocks(c->L), d, F, (const double*)c->dense_pdf, c->ps);
    c->stepped = false;
    return MFLBM_OK;
}
// download into the dense staging (before the first step of a folded domain the staging holds the state as given)
template <class T>
int sp_scatter_t(mflbm_sp_ctx* c) {
    const SDev<T>& d = view<T>(c);
    const Fold F = sp_fold_of(c->p);
    const bool folded = F.y || F.z;
    if (folded && !c->stepped) return MFLBM_OK;
    LAUNCH(c, k_sp_scatter<T>, nblocks(d.nc), d, c->dense_pdf, c->ps);
    LAUNCH(c, (k_sp_links<true, T>), nblocks(d.nf), d, c->dense_pdf, c->ps);
    if (folded) LAUNCH(c, k_fold_dense, fold_blocks(c->L), c->L, F, d.map, d.nf, c->dense_pdf, c->ps, 1);
    return MFLBM_OK;
}
template <class T>
int sp_macro_t(mflbm_sp_ctx* c) {
    const long ps = c->ps;
    const SDev<T>& d = view<T>(c);
    LAUNCH(c, k_sp_macro<T>, nblocks(d.nf), d, c->p.force_z, c->fields, c->fields + ps, c->fields + 2 * ps,
           c->fields + 3 * ps);
    return MFLBM_OK;
}

void sp_free_geometry(mflbm_sp_ctx* c) {
    void* bufs[] = {c->compact, c->map, c->cpos, c->wmask, c->wbase};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    c->compact = nullptr;
    c->map = nullptr;
    c->cpos = nullptr;
    c->wmask = nullptr;
    c->wbase = nullptr;
    c->have_geom = false;
    c->have_state = false;
}

int sp_fold_events(mflbm_sp_ctx* c) {
    if (c->ev_used == 0) return MFLBM_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (size_t b = 0; b + 3 <= c->ev_used; b += 3) {
        float a = 0, s = 0;
        HIPCHK(c, hipEventElapsedTime(&a, c->ev[b], c->ev[b + 1]));
        HIPCHK(c, hipEventElapsedTime(&s, c->ev[b], c->ev[b + 2]));
        c->t_sum[0] += a;
        c->t_sum[2] += s;
        c->t_cnt[0]++;
        c->t_cnt[2]++;
    }
    c->ev_used = 0;
    return MFLBM_OK;
}

}  // namespace

extern "C" const char* mflbm_sp_last_error(const mflbm_sp_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

extern "C" int mflbm_sp_create(const mflbm_sp_params* prm, mflbm_sp_ctx** out) {
    if (!out) return fail(nullptr, MFLBM_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!prm) return fail(nullptr, MFLBM_ERR_ARG, "params is NULL");
    if (prm->nx < 3 || prm->ny < 3 || prm->nz < 3) return fail(nullptr, MFLBM_ERR_ARG, "lattice must be at least 3^3");
    if (prm->wall_x_min != 1 || prm->wall_x_max != 1)
        return fail(nullptr, MFLBM_ERR_ARG, "x needs solid walls at both ends (the reference rejects anything else)");
    if (!(prm->la_nu > 0.0)) return fail(nullptr, MFLBM_ERR_ARG, "viscosity must be positive");
    if (prm->precision < 0 || prm->precision > 2)
        return fail(nullptr, MFLBM_ERR_ARG, "precision must be 0 (FP64), 1 (FP32) or 2 (FP32 storage, FP64 collision), got %d",
                    prm->precision);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(nullptr, MFLBM_ERR_NODEV, "no HIP device visible");
    mflbm_sp_ctx* c = new (std::nothrow) mflbm_sp_ctx();
    if (!c) return fail(nullptr, MFLBM_ERR_NOMEM, "context allocation failed");
    c->p = *prm;
    c->cp = sp_collide_params(prm->la_nu, prm->mrt_preset, prm->force_z);
    c->cp32 = to_float(c->cp);
    c->fp32 = prm->precision != 0;
    c->mixed = prm->precision == 2;
    Layout& L = c->L;
    L = make_layout(prm->nx, prm->ny, prm->nz);
    if (L.total > (1L << 31) - 1) {
        delete c;
        return fail(nullptr, MFLBM_ERR_ARG, "lattice too large for 32-bit node offsets");
    }
    const long ps = ((L.total + 31) / 32) * 32;
    c->ps = ps;
    auto bail = [&](int code) {
        std::string m = c->err;
        mflbm_sp_destroy(c);
        return fail(nullptr, code, "%s", m.c_str());
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(c, MFLBM_ERR_HIP, "hipStreamCreate failed"));
    const long plane = (long)(L.nx + 2) * (L.ny + 2);
    if (hipMalloc(&c->dense_pdf, sizeof(double) * ps * 19) != hipSuccess ||
        hipMalloc(&c->fields, sizeof(double) * ps * 4) != hipSuccess || hipMalloc(&c->walls, (size_t)ps) != hipSuccess ||
        hipMalloc(&c->planes, sizeof(double) * plane * 20) != hipSuccess ||
        hipMalloc(&c->red, sizeof(double) * (size_t)L.nz * 3) != hipSuccess)
        return bail(fail(c, MFLBM_ERR_NOMEM, "device allocation of %.2f GB failed", 23.0 * 8.0 * ps / 1e9));
    if (hipMemset(c->dense_pdf, 0, sizeof(double) * ps * 19) != hipSuccess ||
        hipMemset(c->fields, 0, sizeof(double) * ps * 4) != hipSuccess || hipMemset(c->walls, 1, (size_t)ps) != hipSuccess ||
        hipMemset(c->planes, 0, sizeof(double) * plane * 20) != hipSuccess)
        return bail(fail(c, MFLBM_ERR_HIP, "hipMemset failed"));
    SDev<double>& d = c->d;
    d.L = L;
    d.walls = c->walls;
    d.w_in = c->planes;
    d.cf = c->planes + plane;
    c->d32.L = L;
    c->d32.walls = c->walls;
    c->d32.w_in = d.w_in;
    c->d32.cf = d.cf;
    if (hipDeviceSynchronize() != hipSuccess) return bail(fail(c, MFLBM_ERR_HIP, "hipDeviceSynchronize failed"));
    *out = c;
    return MFLBM_OK;
}

extern "C" void mflbm_sp_destroy(mflbm_sp_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto e : c->ev)
        if (e) (void)hipEventDestroy(e);
    sp_free_geometry(c);
    void* bufs[] = {c->dense_pdf, c->fields, c->walls, c->planes, c->red};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" int mflbm_sp_set_scalars(mflbm_sp_ctx* c, double force_z, double rho_in, double uin_avg, double relaxation) {
    if (!c) return MFLBM_ERR_ARG;
    c->p.force_z = force_z;
    c->p.rho_in = rho_in;
    c->p.uin_avg = uin_avg;
    c->p.relaxation = relaxation;
    c->cp.force_z = force_z;
    c->cp32.force_z = (float)force_z;
    return MFLBM_OK;
}

extern "C" int mflbm_sp_upload_geometry(mflbm_sp_ctx* c, const int8_t* walls_h2, const double* w_in) {
    if (!c) return MFLBM_ERR_ARG;
    if (!walls_h2) return fail(c, MFLBM_ERR_ARG, "walls is NULL");
    const Layout& L = c->L;
    const mflbm_sp_params& p = c->p;
    if (sp_open_z(p) && p.inlet_bc == 1 && !w_in) return fail(c, MFLBM_ERR_ARG, "velocity inlet needs w_in");
    std::vector<int8_t> wp(L.total, 1);
    pack_host(L, 2, walls_h2, wp.data());
    const TileOrder T{tile_rows(), L};
    CompactSet cset;
    build_compact(L, T, wp, sp_open_z(p), p.kper == 1, p.jper == 1, {}, cset, sp_fold_of(p));
    const int nc = (int)cset.cpos.size();
    LinkSet links;
    build_links(L, cset, links);
    const long esz = c->fp32 ? (long)sizeof(float) : (long)sizeof(double);
    long soff[20];
    soff[0] = 0;
    for (int q = 0; q < 19; ++q) {
        const long len = (((long)nc + (q ? links.nlink[lat::OPP[q]] : 0) + 63) / 64) * 64;
        if (len * esz >= (1L << 32))
            return fail(c, MFLBM_ERR_ARG, "population slice of %ld slots exceeds 32-bit byte offsets", len);
        soff[q + 1] = soff[q] + len;
    }
    sp_free_geometry(c);
    HIPCHK(c, hipMalloc(&c->compact, esz * soff[19]));
    HIPCHK(c, hipMalloc(&c->map, sizeof(int) * L.total));
    HIPCHK(c, hipMalloc(&c->cpos, sizeof(int) * ((size_t)nc + 1)));
    HIPCHK(c, hipMalloc(&c->wmask, sizeof(unsigned long long) * links.wmask.size()));
    HIPCHK(c, hipMalloc(&c->wbase, sizeof(int) * links.wbase.size()));
    HIPCHK(c, hipMemset(c->compact, 0, esz * soff[19]));
    HIPCHK(c, hipMemset(c->fields, 0, sizeof(double) * c->ps * 4));
    HIPCHK(c, hipMemcpy(c->map, cset.cidx.data(), sizeof(int) * L.total, hipMemcpyHostToDevice));
    if (nc > 0) HIPCHK(c, hipMemcpy(c->cpos, cset.cpos.data(), sizeof(int) * nc, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->wmask, links.wmask.data(), sizeof(unsigned long long) * links.wmask.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->wbase, links.wbase.data(), sizeof(int) * links.wbase.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->walls, wp.data(), (size_t)L.total, hipMemcpyHostToDevice));
    if (w_in) HIPCHK(c, hipMemcpy(c->d.w_in, w_in, sizeof(double) * (L.nx + 2) * (L.ny + 2), hipMemcpyHostToDevice));
    SDev<double>& d = c->d;
    SDev<float>& d32 = c->d32;
    for (int q = 0; q < 19; ++q) {
        d.S[q] = c->fp32 ? nullptr : (double*)c->compact + soff[q];
        d32.S[q] = c->fp32 ? (float*)c->compact + soff[q] : nullptr;
    }
    d.map = d32.map = c->map;
    d.cpos = d32.cpos = c->cpos;
    d.nf = d32.nf = cset.nfl;
    d.nc = d32.nc = nc;
    d.wmask = d32.wmask = c->wmask;
    d.wbase = d32.wbase = c->wbase;
    HIPCHK(c, hipDeviceSynchronize());
    c->have_geom = true;
    return MFLBM_OK;
}

extern "C" int mflbm_sp_upload_state(mflbm_sp_ctx* c, const double* const* f, const double* f_convec) {
    if (!c) return MFLBM_ERR_ARG;
    if (!c->have_geom) return fail(c, MFLBM_ERR_STATE, "upload the geometry before the state");
    if (!f) return fail(c, MFLBM_ERR_ARG, "missing population arrays");
    for (int q = 0; q < 19; ++q)
        if (!f[q]) return fail(c, MFLBM_ERR_ARG, "population pointer %d is NULL", q);
    const Layout& L = c->L;
    std::vector<double> buf(L.total, 0.0);
    for (int q = 0; q < 19; ++q) {
        pack_host(L, 1, f[q], buf.data());
        HIPCHK(c, hipMemcpy(c->dense_pdf + q * c->ps, buf.data(), sizeof(double) * L.total, hipMemcpyHostToDevice));
    }
    const size_t plane = (size_t)(L.nx + 2) * (L.ny + 2);
    if (f_convec) HIPCHK(c, hipMemcpy(c->d.cf, f_convec, sizeof(double) * plane * 19, hipMemcpyHostToDevice));
    HIPCHK(c, hipDeviceSynchronize());
    const int rc = c->fp32 ? sp_gather_t<float>(c) : sp_gather_t<double>(c);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->have_state = true;
    return MFLBM_OK;
}

// Populations at the MRT collision's equilibrium of prescribed fields, into the dense staging (halo-1 box):
// the second-order equilibrium of mflbm_sp_init_fields (same expressions, same order) shifted to the
// collision's rest state, f_q = (feq_q - w_q rho) + REST_q rho.  fin: u, v, w, rho (dense; has = which are given).
__global__ void k_sp_equilibrium_dense(Layout L, const double* __restrict__ fin, long ps, int has, double* dense) {
    const long nb = (long)(L.nx + 2) * (L.ny + 2) * (L.nz + 2);
    const long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (t >= nb) return;
    const int i = (int)(t % (L.nx + 2)), j = (int)((t / (L.nx + 2)) % (L.ny + 2)), k = (int)(t / ((long)(L.nx + 2) * (L.ny + 2)));
    const int o = L.at(i, j, k);
    const double u = (has & 1) ? fin[o] : 0.0, v = (has & 2) ? fin[ps + o] : 0.0, w = (has & 4) ? fin[2 * ps + o] : 0.0;
    const double rho1 = (has & 8) ? fin[3 * ps + o] : 1.0;
    const double usqrt = u * u + v * v + w * w;
    const double w0e = lat::W0, w1 = lat::W1, w2 = lat::W2;
    auto axis = [&](double c, double cu) { return rho1 * w1 + rho1 * w1 * (3.0 * cu + 4.5 * c * c - 1.5 * usqrt); };
    auto diag = [&](double cu) { return rho1 * w2 + rho1 * w2 * (3.0 * cu + 4.5 * cu * cu - 1.5 * usqrt); };
    double f[19];
    f[0] = rho1 * w0e + rho1 * w0e * (-1.5 * usqrt);
    f[1] = axis(u, u);
    f[2] = axis(u, -u);
    f[3] = axis(v, v);
    f[4] = axis(v, -v);
    f[5] = axis(w, w);
    f[6] = axis(w, -w);
    f[7] = diag(u + v);
    f[8] = diag(-u + v);
    f[9] = diag(u - v);
    f[10] = diag(-u - v);
    f[11] = diag(u + w);
    f[12] = diag(-u + w);
    f[13] = diag(u - w);
    f[14] = diag(-u - w);
    f[15] = diag(v + w);
    f[16] = diag(-v + w);
    f[17] = diag(v - w);
    f[18] = diag(-v - w);
#pragma unroll
    for (int q = 0; q < 19; ++q) dense[q * ps + o] = (f[q] - lat::W[q] * rho1) + rest_weight(q) * rho1;
}
// A slot (y, q) of a node that is not an interior fluid node, facing an interior fluid node x = y + e_q, takes
// x's population opp(q): what x pulls from it at the first (odd) step is the population x sent to the wall.
__global__ void k_sp_bounce_dense(Layout L, const int8_t* __restrict__ walls, long ps, double* dense) {
    const long nb = (long)(L.nx + 2) * (L.ny + 2) * (L.nz + 2);
    const long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (t >= nb) return;
    const int i = (int)(t % (L.nx + 2)), j = (int)((t / (L.nx + 2)) % (L.ny + 2)), k = (int)(t / ((long)(L.nx + 2) * (L.ny + 2)));
    auto fluid = [&](int a, int b, int cz) {
        return a >= 1 && a <= L.nx && b >= 1 && b <= L.ny && cz >= 1 && cz <= L.nz && walls[L.at(a, b, cz)] == 0;
    };
    if (fluid(i, j, k)) return;
    const int o = L.at(i, j, k);
    for (int q = 1; q < 19; ++q) {
        const int a = i + lat::EX[q], b = j + lat::EY[q], cz = k + lat::EZ[q];
        if (fluid(a, b, cz)) dense[q * ps + o] = dense[lat::OPP[q] * ps + L.at(a, b, cz)];
    }
}

extern "C" int mflbm_sp_set_equilibrium(mflbm_sp_ctx* c, const double* rho, const double* u, const double* v,
                                        const double* w) {
    if (!c) return MFLBM_ERR_ARG;
    if (!c->have_geom) return fail(c, MFLBM_ERR_STATE, "upload the geometry before the state");
    const Layout& L = c->L;
    const long ps = c->ps;
    // the macroscopic-field buffers are free between compute_macro calls: they take the prescribed fields
    std::vector<double> buf(L.total, 0.0);
    const double* src[4] = {u, v, w, rho};
    int has = 0;
    for (int m = 0; m < 4; ++m) {
        if (!src[m]) continue;
        has |= 1 << m;
        pack_host(L, 1, src[m], buf.data());
        HIPCHK(c, hipMemcpy(c->fields + m * ps, buf.data(), sizeof(double) * L.total, hipMemcpyHostToDevice));
    }
    const long nb = (long)(L.nx + 2) * (L.ny + 2) * (L.nz + 2);
    LAUNCH(c, k_sp_equilibrium_dense, nblocks(nb), L, (const double*)c->fields, ps, has, c->dense_pdf);
    LAUNCH(c, k_sp_bounce_dense, nblocks(nb), L, (const int8_t*)c->walls, ps, c->dense_pdf);
    const int rc = c->fp32 ? sp_gather_t<float>(c) : sp_gather_t<double>(c);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->have_state = true;
    return MFLBM_OK;
}

extern "C" int mflbm_sp_download_state(mflbm_sp_ctx* c, double* const* f, double* f_convec) {
    if (!c) return MFLBM_ERR_ARG;
    if (!c->have_state) return fail(c, MFLBM_ERR_STATE, "no state uploaded");
    const Layout& L = c->L;
    const int rc = c->fp32 ? sp_scatter_t<float>(c) : sp_scatter_t<double>(c);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<double> buf(L.total);
    for (int q = 0; q < 19 && f; ++q) {
        if (!f[q]) continue;
        HIPCHK(c, hipMemcpy(buf.data(), c->dense_pdf + q * c->ps, sizeof(double) * L.total, hipMemcpyDeviceToHost));
        unpack_host(L, 1, buf.data(), f[q]);
    }
    const size_t plane = (size_t)(L.nx + 2) * (L.ny + 2);
    if (f_convec) HIPCHK(c, hipMemcpy(f_convec, c->d.cf, sizeof(double) * plane * 19, hipMemcpyDeviceToHost));
    return MFLBM_OK;
}

extern "C" int mflbm_sp_step(mflbm_sp_ctx* c, int32_t ntime) {
    if (!c) return MFLBM_ERR_ARG;
    if (!c->have_geom || !c->have_state) return fail(c, MFLBM_ERR_STATE, "upload geometry and state first");
    return sp_launch_step(c, ntime);
}

extern "C" int mflbm_sp_run(mflbm_sp_ctx* c, int32_t ntime0, int32_t nsteps) {
    if (!c) return MFLBM_ERR_ARG;
    if (!c->have_geom || !c->have_state) return fail(c, MFLBM_ERR_STATE, "upload geometry and state first");
    if (nsteps < 0) return fail(c, MFLBM_ERR_ARG, "nsteps < 0");
    for (int s = 0; s < nsteps; ++s) {
        const int rc = sp_launch_step(c, ntime0 + s);
        if (rc) return rc;
    }
    return MFLBM_OK;
}

extern "C" int mflbm_sp_compute_macro(mflbm_sp_ctx* c) {
    if (!c) return MFLBM_ERR_ARG;
    if (!c->have_state) return fail(c, MFLBM_ERR_STATE, "no state uploaded");
    return c->fp32 ? sp_macro_t<float>(c) : sp_macro_t<double>(c);
}

extern "C" int mflbm_sp_monitor(mflbm_sp_ctx* c, double* layers, double* umax2) {
    if (!c) return MFLBM_ERR_ARG;
    const int rc = mflbm_sp_compute_macro(c);
    if (rc) return rc;
    const Layout& L = c->L;
    const long ps = c->ps;
    LAUNCH(c, k_sp_monitor_layers, (unsigned)L.nz, L, c->fields, c->fields + ps, c->fields + 2 * ps, c->fields + 3 * ps,
           c->red);
    std::vector<double> h((size_t)L.nz * 3);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->red, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (layers) std::memcpy(layers, h.data(), sizeof(double) * 2 * L.nz);
    if (umax2) {
        double um = 0.0;
        for (int k = 0; k < L.nz; ++k)
            if (um < h[2 * (size_t)L.nz + k]) um = h[2 * (size_t)L.nz + k];
        *umax2 = um;
    }
    return MFLBM_OK;
}

extern "C" int mflbm_sp_download_macro(mflbm_sp_ctx* c, double* u, double* v, double* w, double* rho) {
    if (!c) return MFLBM_ERR_ARG;
    const Layout& L = c->L;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<double> buf(L.total);
    double* dst[4] = {u, v, w, rho};
    for (int m = 0; m < 4; ++m) {
        if (!dst[m]) continue;
        HIPCHK(c, hipMemcpy(buf.data(), c->fields + m * c->ps, sizeof(double) * L.total, hipMemcpyDeviceToHost));
        unpack_host(L, 1, buf.data(), dst[m]);
    }
    return MFLBM_OK;
}

extern "C" int mflbm_sp_synchronize(mflbm_sp_ctx* c) {
    if (!c) return MFLBM_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MFLBM_OK;
}

extern "C" int mflbm_sp_timing_enable(mflbm_sp_ctx* c, int32_t enable) {
    if (!c) return MFLBM_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->timing = enable != 0;
    c->ev_used = 0;
    for (int m = 0; m < 3; ++m) {
        c->t_sum[m] = 0;
        c->t_cnt[m] = 0;
    }
    return MFLBM_OK;
}

extern "C" int mflbm_sp_timing(mflbm_sp_ctx* c, int32_t which, double* mean_ms, int64_t* count) {
    if (!c || (which != 0 && which != 2)) return MFLBM_ERR_ARG;
    const int rc = sp_fold_events(c);
    if (rc) return rc;
    if (mean_ms) *mean_ms = c->t_cnt[which] ? c->t_sum[which] / c->t_cnt[which] : 0.0;
    if (count) *count = c->t_cnt[which];
    return MFLBM_OK;
}
