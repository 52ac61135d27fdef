// gfx950 implementation of MF-LBM's per-time-step lattice update (two-colour D3Q19 MRT, FP64) and its
// C ABI.
//
// Storage (device_common.h): dense fields (phi, interface normal, curvature, walls, macroscopic fields) on
// one padded grid; the 38 population arrays (f0..f18 fluid 1, g0..g18 fluid 2) compact over the interior
// fluid nodes plus the support nodes the update, the boundary conditions and the periodic exchange
// address.  A slot is therefore stored exactly where the reference keeps a value that is ever read, so the
// update is the reference's, slot for slot, while HBM traffic follows the fluid nodes instead of whole
// cache lines of a mostly solid grid.  A dense copy of all 38 arrays is kept only as the staging area of
// upload/download.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "collide.h"
#include "device_common.h"
#include "lattice.h"
#include "mflbm.h"

using namespace mflbm_dev;

namespace {

// Population slots: slice q holds, per slot, the pair (fluid 1, fluid 2) of direction q (FP32 storage: (s, a)), so
// p[q] is the base of the fluid-1 halves and p[19 + q] = p[q] + one element the base of the fluid-2 halves; slot i
// of either colour is element 2 i from its base.
struct SlotTable {
    double* p[38];
};

// device view of the state passed to kernels by value
struct Dev {
    Layout L;
    // Population slice q holds the slots of the compact nodes at [0, nc) and, from index nc, the link slots
    // of direction r = opp(q): for a fluid node x whose neighbour x + e_r is not a fluid node, the slot
    // (x + e_r, opp(r)) is read and written only by x (the AA update pulls it as the incoming population
    // opp(r) and pushes the outgoing population r into it), so it is kept in x's order instead of at the
    // neighbour's compact index.  Link slots of direction r are packed in fluid-node order: a wave of 64
    // consecutive fluid nodes finds its links of one direction contiguous (index nc + wbase + rank in wave).
    SlotTable S;
    const int* map;   // padded offset -> compact index, -1 where nothing is stored
    const int* pmap;  // padded offset -> compact index of an interior fluid node, -1 elsewhere (a folded halo node
                      // maps to its image in map but not here: its phi is the dense one the periodic copy writes)
    const int* cpos;  // compact index -> padded offset
    int nf;           // interior fluid nodes
    int nc;           // all compact nodes
    const unsigned long long* wmask;   // per wave and direction: fluid nodes with a link (device_common.h)
    const int* wbase;                  // per wave and direction: links before the wave
    double* force;    // per fluid node: cn_x, cn_y, cn_z, 0.5*gamma*curv*c_norm (stride fs)
    long fs;
    double* phi;
    // phi at the fluid nodes by compact index, as the lattice update writes it (whole cache lines).  The chain and
    // the boundary conditions read it through phi_now; the dense phi takes these values before any other reader
    // (k_phi_dense)
    double* phic;
    // The normal field, one 32-byte record (cn_x, cn_y, cn_z, c_norm) per dense node: a neighbour's normal is one
    // cache-line access in the stencils that read it (normal extrapolation, curvature)
    double4* cn;
    // FP32 storage modes: the normal field in FP32 (16-byte records), the same dense layout
    float4* cn32;
    double* curv;     // dense curvature, filled only for diagnostics
    // Brick flags of the list chain.  The padded grid is cut into bricks of 64 nodes (brick_of).  Away
    // from the interface phi is +-1 to within rounding, and there the ISO4 gradient is below the 1e-6 at which
    // the reference zeroes the normal (Phase_gradient.F90:63-79): if every phi a node's normal depends on (its
    // stencil, and the fluid neighbours of the solid nodes in it) lies within PHI_BAND of the same +-1, the
    // normal is exactly +0 whatever the values, and so is every normal extrapolated from such normals.
    //   notm[b], notp[b]: a fluid node of brick b holds a phi farther than PHI_BAND from -1 (from +1), or b is
    //            static; set by the lattice update of the current step, read and reset to bstatic by the next
    //            chain's plan (two buffers in turn)
    //   bstatic[b]: brick b holds a non-solid node outside the interior (halo planes that boundary conditions,
    //            the periodic exchange or nobody write) or an interior node within 4 of a periodic face (whose
    //            phi the periodic exchange copies into the halo): never quiet
    //   nzb[b]:  a stored normal of brick b has a bit set
    //   nzd[b]:  nzb of brick b or one of its 26 neighbours is set: the lattice update reads the stencil
    // From them each chain's plan (k_brick_plan) derives, per brick, q1b (every brick within one of b is quiet with
    // one sign: the normals of b are +0, not evaluated) and q2b (q1b of every brick within one of b: the
    // extrapolated normals of b are +0, and no evaluated normal reads the extrapolated phi of b's solid nodes; not
    // evaluated, downloads refresh them), and zero-fills skipped normals that had a bit set (nzb).
    unsigned char* notm;
    unsigned char* notp;
    const unsigned char* bstatic;
    unsigned char* nzb;
    unsigned char* nzd;
    int nbx, nby, nbz;  // bricks per axis of the padded grid
    const int8_t* walls;
    double* w_in;     // (nx+2)(ny+2), reference layout
    double* cf;       // convective-outlet history f: 19 planes (nx+2)(ny+2)
    double* cg;
    double* cphi;     // (nx+2)(ny+2)
};

__device__ __forceinline__ int link_index(const Dev& d, int x, int r) { return link_slot(d.nc, d.wbase, d.wmask, x, r); }

// index of slot (direction q, padded offset o) in slice q through the compact map.  Slot q of a support node o
// faces the node o + e_q; when that node is fluid the slot is its link slot of direction opp(q).
__device__ __forceinline__ int slot_index(const Dev& d, int q, int o) {
    const int c = d.map[o];
    if (c >= d.nf) {
        const int x = d.map[o + d.L.off(q)];
        if (x >= 0 && x < d.nf) return link_index(d, x, lat::OPP[q]);
    }
    return c;
}

// Population accessors by (direction, padded offset) for the boundary, exchange and staging code.
// T = double: the slices hold f (slices 0..18) and g (19..37), the reference arithmetic.
// T = float: the slices hold the total deviation s = f + g - REST (0..18) and the colour difference a = f - g
// (19..37) (collide.h); a colour's population is read and written in FP64 through CRef, which rounds s and a
// once per write and leaves the other colour unchanged.
template <class T>
struct Pop;
template <>
struct Pop<double> {
    const Dev* d;
    int c0;   // 0 fluid 1, 19 fluid 2
    __device__ __forceinline__ double& operator()(int q, int o) const { return d->S.p[c0 + q][2 * slot_index(*d, q, o)]; }
    // the slot of direction q with a slot index resolved before (slot_index)
    __device__ __forceinline__ double& at(int q, int i) const { return d->S.p[c0 + q][2 * i]; }
};
struct CRef {
    float* s;
    float* a;
    int q;
    int g;   // 0 fluid 1, 1 fluid 2
    __device__ __forceinline__ operator double() const {
        const double F = (double)*s + rest_weight(q), A = (double)*a;
        return g ? 0.5 * (F - A) : 0.5 * (F + A);
    }
    __device__ __forceinline__ CRef& operator=(double v) {
        const double F = (double)*s + rest_weight(q), A = (double)*a;
        const double other = g ? 0.5 * (F + A) : 0.5 * (F - A);
        *s = (float)(v + other - rest_weight(q));
        *a = (float)(g ? other - v : v - other);
        return *this;
    }
    __device__ __forceinline__ CRef& operator=(const CRef& o) { return *this = (double)o; }
};
template <>
struct Pop<float> {
    const Dev* d;
    int c0;
    __device__ __forceinline__ CRef operator()(int q, int o) const { return at(q, slot_index(*d, q, o)); }
    __device__ __forceinline__ CRef at(int q, int i) const {
        return CRef{(float*)d->S.p[q] + 2 * i, (float*)d->S.p[19 + q] + 2 * i, q, c0 != 0};
    }
};
template <class T>
__device__ __forceinline__ Pop<T> fpop(const Dev& d) { return Pop<T>{&d, 0}; }
template <class T>
__device__ __forceinline__ Pop<T> gpop(const Dev& d) { return Pop<T>{&d, 19}; }
__device__ __forceinline__ int plane_at(const Dev& d, int i, int j) { return i + j * (d.L.nx + 2); }

// brick of padded offset n: 2^BRICK_SX x 2^BRICK_SY x 2^BRICK_SZ padded nodes (16 x 2 x 2: a brick's rows are
// whole 128-byte lines of a dense FP64 field, so the chain stages that take one brick per wave touch few lines
// per access).  The brick flags need every brick at least 2 nodes wide in each axis (a normal depends on phi
// within 2 nodes: its stencil and the fluid neighbours of the stencil's solid nodes).
constexpr int BRICK_SX = 4, BRICK_SY = 1, BRICK_SZ = 1;
static_assert(BRICK_SX >= 1 && BRICK_SY >= 1 && BRICK_SZ >= 1 && BRICK_SX + BRICK_SY + BRICK_SZ == 6,
              "64-node bricks at least 2 nodes wide");
__device__ __forceinline__ int brick_of(const Dev& d, int n) {
    const int kp = n / d.L.pxy;
    const int r = n - kp * d.L.pxy;
    const int jp = r / d.L.px;
    const int ip = r - jp * d.L.px;
    return (ip >> BRICK_SX) + d.nbx * ((jp >> BRICK_SY) + d.nby * (kp >> BRICK_SZ));
}
// Half-width of the band around +-1 within which phi counts as pure.  With every value of an ISO4 stencil within
// eps of one constant, each difference is at most 2 eps and exact (Sterbenz), so |d/dx| <= (1/6 + 8/12) 2 eps and
// |grad phi| <= sqrt(3) (5/3) eps = 2.89 eps (plus relative rounding of order 1e-15); a solid node's extrapolated
// phi is a weighted mean of fluid values within the band, so inside it up to one rounding.  eps = 3e-7 keeps
// |grad phi| below 8.7e-7 < 1e-6.
constexpr double PHI_BAND = 3e-7;
__device__ __forceinline__ void flag_band(const Dev& d, int b, double v) {
    if (!(fabs(v + 1.0) <= PHI_BAND)) d.notm[b] = 1;
    if (!(fabs(v - 1.0) <= PHI_BAND)) d.notp[b] = 1;
}
// current phi of padded node n: the compact value at a fluid node, the dense value elsewhere (one load from the
// selected address)
__device__ __forceinline__ double phi_now(const Dev& d, int n) {
    const int c = d.pmap[n];
    return *((unsigned)c < (unsigned)d.nf ? d.phic + c : d.phi + n);
}
// phi_now around node n, for the stencils (DENSE: the dense phi, where it is current at every node)
template <bool DENSE = false>
__device__ __forceinline__ double phi_at(const Dev& d, int n) {
    return DENSE ? d.phi[n] : phi_now(d, n);
}
template <bool DENSE = false>
struct PhiNowAt {
    const Dev* d;
    int n;
    __device__ __forceinline__ double operator()(int a, int b, int c) const {
        return phi_at<DENSE>(*d, n + a + b * d->L.px + c * d->L.pxy);
    }
};


// ------------------------------------------------------------------------------------------------------
// lattice update: collision + recolouring + AA-pattern streaming of both colours, one thread per fluid node
// ------------------------------------------------------------------------------------------------------

// Population slots.  The 38 slice base pointers travel in the kernel arguments, so each is a scalar value
// the compiler cannot relate to the others: every access becomes one global load/store with a scalar base
// and a 32-bit vector byte offset, instead of a 64-bit vector address per slot kept live across the collision.
// upload_geometry guarantees that every slice is below 2^32 bytes.
// The two colours of a slot are adjacent (f_q, g_q or s_q, a_q): one 16-byte (FP64) or 8-byte (FP32) access
// moves both, which halves the scattered accesses of the odd step.
__device__ __forceinline__ double2 ld_pair(const double* slice, unsigned m) {
    return *(const double2*)((const char*)slice + m * 16u);
}
__device__ __forceinline__ void st_pair(double* slice, unsigned m, double f, double g) {
    *(double2*)((char*)slice + m * 16u) = double2{f, g};
}
__device__ __forceinline__ float2 ld_pair32(const double* slice, unsigned m) {
    return *(const float2*)((const char*)slice + m * 8u);
}
__device__ __forceinline__ void st_pair32(double* slice, unsigned m, float s, float a) {
    *(float2*)((char*)slice + m * 8u) = float2{s, a};
}
// One slot (direction q, index i) of both colours held in registers.  The velocity inlet and the convective outlet
// load every slot they touch before their first store (a load after a store through another pointer of the same
// type must wait for it, and the slices cannot be told apart at compile time), update the registers in the order of
// the colour-by-colour code, and store each slot once.  get/set mirror Pop<T>::at: FP64 holds (f, g); FP32 holds
// (s, a) and a colour's value is read and written in FP64 exactly as CRef does.
template <class T>
struct SlotReg;
template <>
struct SlotReg<double> {
    double v[2];
    __device__ __forceinline__ void load(const Dev& d, int q, int i) {
        const double2 p = *(const double2*)(d.S.p[q] + 2 * (long)i);
        v[0] = p.x;
        v[1] = p.y;
    }
    __device__ __forceinline__ void store(const Dev& d, int q, int i) const {
        *(double2*)(d.S.p[q] + 2 * (long)i) = double2{v[0], v[1]};
    }
    __device__ __forceinline__ double get(int c, int) const { return v[c]; }
    __device__ __forceinline__ void set(int c, int, double x) { v[c] = x; }
};
template <>
struct SlotReg<float> {
    float s, a;
    __device__ __forceinline__ void load(const Dev& d, int q, int i) {
        const float2 p = *(const float2*)((const float*)d.S.p[q] + 2 * (long)i);
        s = p.x;
        a = p.y;
    }
    __device__ __forceinline__ void store(const Dev& d, int q, int i) const {
        *(float2*)((float*)d.S.p[q] + 2 * (long)i) = float2{s, a};
    }
    __device__ __forceinline__ double get(int c, int q) const {
        const double F = (double)s + rest_weight(q), A = (double)a;
        return c ? 0.5 * (F - A) : 0.5 * (F + A);
    }
    __device__ __forceinline__ void set(int c, int q, double x) {
        const double F = (double)s + rest_weight(q), A = (double)a;
        const double other = c ? 0.5 * (F + A) : 0.5 * (F - A);
        s = (float)(x + other - rest_weight(q));
        a = (float)(c ? other - x : x - other);
    }
};
// A slot of the velocity inlet or the convective outlet.  HOIST: held in registers (SlotReg); otherwise read and
// written in place through Pop<T> in the same order, with fewer registers: the conditions that run in extra
// work-groups of the chain's plan launch would otherwise lower the occupancy of the whole launch.
template <class T, bool HOIST>
struct FaceSlot;
template <class T>
struct FaceSlot<T, true> : SlotReg<T> {};
template <class T>
struct FaceSlot<T, false> {
    const Dev* d;
    int i;
    __device__ __forceinline__ void load(const Dev& dd, int, int ii) {
        d = &dd;
        i = ii;
    }
    __device__ __forceinline__ void store(const Dev&, int, int) const {}
    __device__ __forceinline__ double get(int c, int q) const { return (double)Pop<T>{d, c ? 19 : 0}.at(q, i); }
    __device__ __forceinline__ void set(int c, int q, double x) const { Pop<T>{d, c ? 19 : 0}.at(q, i) = x; }
};

template <class NT>
__device__ __forceinline__ double curvature_at(const Dev& d, int n);
// the curvature where every normal of the stencil is +0
__device__ __forceinline__ double quiet_curvature();

// ODD = false: kernel_even_color (Kernel_multiphase.F90:371-725): read swapped own slots, write own slots.
// ODD = true : kernel_odd_color  (Kernel_multiphase.F90:6-362): pull f_q(x - e_q), push to f_opp(q)(x + e_q).
// The CSF force input (curvature, Phase_gradient.F90:112-201, and 0.5*gamma*curv*c_norm,
// Kernel_multiphase.F90:116) is either evaluated here from the dense normal field the list kernels of the chain
// left (STENCIL), or read from the compact force arrays the fused chain (k_chain_force) wrote.
template <bool ODD, bool STENCIL>
__global__ void __launch_bounds__(BLOCK)
__attribute__((amdgpu_waves_per_eu(2, 8)))
k_collide(Dev d, CollideParams P, int per) {
    const SlotTable& S = d.S;
    const int m = (int)blockIdx.x * BLOCK + (int)threadIdx.x;
    if (m >= d.nf) return;
    const int nn = d.cpos[m];
    double cnx, cny, cnz, tf;
    if (STENCIL) {
        if (d.nzd[brick_of(d, nn)]) {
            const double4 c4 = d.cn[nn];
            cnx = c4.x;
            cny = c4.y;
            cnz = c4.z;
            tf = 0.5 * P.gamma * curvature_at<double>(d, nn) * c4.w;
        } else {   // every normal of the stencil is +0: the same expressions on those values
            cnx = 0.0;
            cny = 0.0;
            cnz = 0.0;
            tf = 0.5 * P.gamma * quiet_curvature() * 0.0;
        }
        // keep the stencil phase and the population phase apart so their registers are not live together
        __builtin_amdgcn_sched_barrier(0);
    } else {
        cnx = d.force[m];
        cny = d.force[d.fs + m];
        cnz = d.force[2 * d.fs + m];
        tf = d.force[3 * d.fs + m];
    }
    double a[19], b[19];
    unsigned nb[19];   // odd: per direction r, the index of the slot (x + e_r, opp(r)) in slice opp(r)
    {
        const double2 v = ld_pair(S.p[0], m);
        a[0] = v.x;
        b[0] = v.y;
    }
    if (ODD) {
        odd_neighbour_slots<false>(d.map, d.L, d.nc, d.wbase, d.wmask, m, nn, nb);
#pragma unroll
        for (int q = 1; q < 19; ++q) {   // slot q of the upwind node x - e_q = x + e_opp(q)
            const double2 v = ld_pair(S.p[q], nb[lat::OPP[q]]);
            a[q] = v.x;
            b[q] = v.y;
        }
    } else {
#pragma unroll
        for (int q = 1; q < 19; ++q) {
            const double2 v = ld_pair(S.p[lat::OPP[q]], m);
            a[q] = v.x;
            b[q] = v.y;
        }
    }
    const double phi = collide_node_tf(a, b, cnx, cny, cnz, tf, P);
    flag_band(d, brick_of(d, nn), phi);
    d.phic[m] = phi;
    st_pair(S.p[0], m, a[0], b[0]);
#pragma unroll
    for (int q = 1; q < 19; ++q) {   // odd: slot opp(q) of x + e_q
        const int qs = ODD ? lat::OPP[q] : q;
        const unsigned ms = ODD ? nb[q] : (unsigned)m;
        st_pair(S.p[qs], ms, a[q], b[q]);
    }
}

// FP32 storage (s = f + g - REST, a = f - g, collide.h): the same update, the collision and recolouring in C
// (float, or double for the mixed mode).  The normal, curvature and phi stay FP64 (the colour-gradient chain is
// shared with the FP64 path).
template <bool ODD, class C, bool STENCIL>
__global__ void __launch_bounds__(BLOCK) k_collide32(Dev d, CollideParamsDev<C> P, double gamma, int per) {
    const SlotTable& S = d.S;
    const int m = (int)blockIdx.x * BLOCK + (int)threadIdx.x;
    if (m >= d.nf) return;
    const int nn = d.cpos[m];
    C cnx, cny, cnz, tf;
    if (STENCIL) {
        if (d.nzd[brick_of(d, nn)]) {
            const float4 c4 = d.cn32[nn];
            cnx = (C)c4.x;
            cny = (C)c4.y;
            cnz = (C)c4.z;
            tf = (C)(0.5 * gamma * curvature_at<float>(d, nn) * (double)c4.w);
        } else {
            cnx = (C)0.0f;
            cny = (C)0.0f;
            cnz = (C)0.0f;
            tf = (C)(0.5 * gamma * quiet_curvature() * (double)0.0f);
        }
        __builtin_amdgcn_sched_barrier(0);
    } else {   // the force arrays hold the FP32 normal widened to FP64
        cnx = (C)d.force[m];
        cny = (C)d.force[d.fs + m];
        cnz = (C)d.force[2 * d.fs + m];
        tf = (C)d.force[3 * d.fs + m];
    }
    C s[19], a[19];
    unsigned nb[19];
    {
        const float2 v = ld_pair32(S.p[0], (unsigned)m);
        s[0] = v.x;
        a[0] = v.y;
    }
    if (ODD) {
        odd_neighbour_slots<false>(d.map, d.L, d.nc, d.wbase, d.wmask, m, nn, nb);
#pragma unroll
        for (int q = 1; q < 19; ++q) {
            const float2 v = ld_pair32(S.p[q], nb[lat::OPP[q]]);
            s[q] = v.x;
            a[q] = v.y;
        }
    } else {
#pragma unroll
        for (int q = 1; q < 19; ++q) {
            const float2 v = ld_pair32(S.p[lat::OPP[q]], (unsigned)m);
            s[q] = v.x;
            a[q] = v.y;
        }
    }
    const C phi = collide_node_dev<C>(s, a, cnx, cny, cnz, tf, P);
    flag_band(d, brick_of(d, nn), (double)phi);
    d.phic[m] = (double)phi;
    st_pair32(S.p[0], (unsigned)m, (float)s[0], (float)a[0]);
#pragma unroll
    for (int q = 1; q < 19; ++q) {
        const int qs = ODD ? lat::OPP[q] : q;
        const unsigned ms = ODD ? nb[q] : (unsigned)m;
        st_pair32(S.p[qs], ms, (float)s[q], (float)a[q]);
    }
}

// slot i of direction q <-> (f, g) of the dense staging arrays at padded offset o
template <class T>
__device__ __forceinline__ void slot_from_dense(const Dev& d, int q, int i, const double* dense, long ps, long o) {
    if (sizeof(T) == 8) {
        d.S.p[q][2 * i] = dense[q * ps + o];
        d.S.p[19 + q][2 * i] = dense[(19 + q) * ps + o];
    } else {
        const double f = dense[q * ps + o], g = dense[(19 + q) * ps + o];
        ((float*)d.S.p[q])[2 * i] = (float)(f + g - rest_weight(q));
        ((float*)d.S.p[19 + q])[2 * i] = (float)(f - g);
    }
}
template <class T>
__device__ __forceinline__ void slot_to_dense(const Dev& d, int q, int i, double* dense, long ps, long o) {
    if (sizeof(T) == 8) {
        dense[q * ps + o] = d.S.p[q][2 * i];
        dense[(19 + q) * ps + o] = d.S.p[19 + q][2 * i];
    } else {
        const double F = (double)((const float*)d.S.p[q])[2 * i] + rest_weight(q), A = (double)((const float*)d.S.p[19 + q])[2 * i];
        dense[q * ps + o] = 0.5 * (F + A);
        dense[(19 + q) * ps + o] = 0.5 * (F - A);
    }
}
// dense staging <-> compact storage
template <class T>
__global__ void k_gather(Dev d, const double* __restrict__ dense, long ps) {
    const long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (t >= d.nc) return;
    const int n = d.cpos[t];
    for (int q = 0; q < 19; ++q) slot_from_dense<T>(d, q, (int)t, dense, ps, n);
}
template <class T>
__global__ void k_scatter(Dev d, double* __restrict__ dense, long ps) {
    const long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (t >= d.nc) return;
    const int n = d.cpos[t];
    for (int q = 0; q < 19; ++q) slot_to_dense<T>(d, q, (int)t, dense, ps, n);
}
// link slots <-> dense staging (after k_gather / after k_scatter: a link slot holds the current value)
template <bool TO_DENSE, class T>
__global__ void k_links(Dev d, double* __restrict__ dense, long ps) {
    const int m = blockIdx.x * BLOCK + threadIdx.x;
    if (m >= d.nf) return;
    const int n = d.cpos[m];
    for (int r = 1; r < 19; ++r) {
        const int i = link_index(d, m, r);
        if (i < 0) continue;
        const long o = n + d.L.off(r);
        const int q = lat::OPP[r];
        if (TO_DENSE) slot_to_dense<T>(d, q, i, dense, ps, o);
        else slot_from_dense<T>(d, q, i, dense, ps, o);
    }
}

// Upload of a folded domain: the exchanged slot (q, y) of each image takes the value of its halo slot (q, h).  A
// state is uploaded as the state before an odd step: that step pulls (q, h), and the push exchange after it leaves
// (q, y) equal to (q, h) (every reference run starts at step 1; a checkpoint written after either exchange already
// holds equal values).  The dense staging keeps the state as given.
template <class T>
__global__ void k_fold_upload(Dev d, Fold F, const double* __restrict__ dense, long ps) {
    fold_exchanged(d.L, F, (long)blockIdx.x * BLOCK + threadIdx.x, [&](int h, int y, int q) {
        if (d.map[y] < 0) return;
        const Pop<T> f = fpop<T>(d), g = gpop<T>(d);
        f(q, y) = dense[(long)q * ps + h];
        g(q, y) = dense[(long)(19 + q) * ps + h];
    });
}

// ------------------------------------------------------------------------------------------------------
// periodic wrap of populations and phi (one rank of the reference's Cartesian MPI exchange, Mpi.F90)
// ------------------------------------------------------------------------------------------------------

// directions with e_z = +1 / -1 and e_y = +1 / -1 (the five populations per colour a face exchanges)
__constant__ int ZP[5] = {5, 11, 12, 15, 16};
__constant__ int ZM[5] = {6, 14, 13, 18, 17};
__constant__ int YP[5] = {3, 7, 8, 15, 17};
__constant__ int YM[5] = {4, 10, 9, 16, 18};

// the stored values of both slices move unchanged (same direction on both sides)
template <class T>
__device__ __forceinline__ void copy_both(const Dev& d, int q, int dst, int src) {
    T* f = (T*)d.S.p[q];
    T* g = (T*)d.S.p[19 + q];
    const int a = slot_index(d, q, dst), b = slot_index(d, q, src);
    f[2 * a] = f[2 * b];
    g[2 * a] = g[2 * b];
}
// after an even step: outgoing boundary-layer populations into the opposite halo
// (mpi_pdf_pull_async_pack / _sync_update, Mpi.F90:101-346)
template <class T>
__global__ void k_wrap_pull_z(Dev d) {
    const Layout L = d.L;
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= L.nx * L.ny) return;
    const int i = t % L.nx + 1, j = t / L.nx + 1;
    for (int m = 0; m < 5; ++m) {
        copy_both<T>(d, ZP[m], L.at(i, j, 0), L.at(i, j, L.nz));
        copy_both<T>(d, ZM[m], L.at(i, j, L.nz + 1), L.at(i, j, 1));
    }
}
template <class T>
__global__ void k_wrap_pull_y(Dev d) {
    const Layout L = d.L;
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= L.nx * L.nz) return;
    const int i = t % L.nx + 1, k = t / L.nx + 1;
    for (int m = 0; m < 5; ++m) {
        copy_both<T>(d, YP[m], L.at(i, 0, k), L.at(i, L.ny, k));
        copy_both<T>(d, YM[m], L.at(i, L.ny + 1, k), L.at(i, 1, k));
    }
}
template <class T>
__global__ void k_wrap_pull_edge(Dev d) {
    const Layout L = d.L;
    const int i = blockIdx.x * BLOCK + threadIdx.x + 1;
    if (i > L.nx) return;
    copy_both<T>(d, 18, L.at(i, L.ny + 1, L.nz + 1), L.at(i, 1, 1));
    copy_both<T>(d, 16, L.at(i, L.ny + 1, 0), L.at(i, 1, L.nz));
    copy_both<T>(d, 17, L.at(i, 0, L.nz + 1), L.at(i, L.ny, 1));
    copy_both<T>(d, 15, L.at(i, 0, 0), L.at(i, L.ny, L.nz));
}
// after an odd step: populations pushed into the halo go back to the periodic image
// (mpi_pdf_push_async_pack / _sync_update, Mpi.F90:354-598); faces first, x-parallel edges last
template <class T>
__global__ void k_wrap_push_z(Dev d) {
    const Layout L = d.L;
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= L.nx * L.ny) return;
    const int i = t % L.nx + 1, j = t / L.nx + 1;
    for (int m = 0; m < 5; ++m) {
        copy_both<T>(d, ZM[m], L.at(i, j, 1), L.at(i, j, L.nz + 1));
        copy_both<T>(d, ZP[m], L.at(i, j, L.nz), L.at(i, j, 0));
    }
}
template <class T>
__global__ void k_wrap_push_y(Dev d) {
    const Layout L = d.L;
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= L.nx * L.nz) return;
    const int i = t % L.nx + 1, k = t / L.nx + 1;
    for (int m = 0; m < 5; ++m) {
        copy_both<T>(d, YM[m], L.at(i, 1, k), L.at(i, L.ny + 1, k));
        copy_both<T>(d, YP[m], L.at(i, L.ny, k), L.at(i, 0, k));
    }
}
template <class T>
__global__ void k_wrap_push_edge(Dev d) {
    const Layout L = d.L;
    const int i = blockIdx.x * BLOCK + threadIdx.x + 1;
    if (i > L.nx) return;
    copy_both<T>(d, 15, L.at(i, L.ny, L.nz), L.at(i, 0, 0));
    copy_both<T>(d, 17, L.at(i, L.ny, 1), L.at(i, 0, L.nz + 1));
    copy_both<T>(d, 16, L.at(i, 1, L.nz), L.at(i, L.ny + 1, 0));
    copy_both<T>(d, 18, L.at(i, 1, 1), L.at(i, L.ny + 1, L.nz + 1));
}
// phi halo of 4 layers in the periodic directions (MPI_phi_async_pack / _sync_update, Mpi.F90:608-867):
// each halo node of the y-z plane (x interior) takes the value of its periodic image
__global__ void k_wrap_phi(Dev d, int wrap_y, int wrap_z) {
    const Layout L = d.L;
    const int ny8 = L.ny + 8, nz8 = L.nz + 8;
    const long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (t >= (long)L.nx * ny8 * nz8) return;
    const int i = (int)(t % L.nx) + 1;
    const int j = (int)((t / L.nx) % ny8) - 3;
    const int k = (int)(t / ((long)L.nx * ny8)) - 3;
    const bool gy = j < 1 || j > L.ny, gz = k < 1 || k > L.nz;
    if (!gy && !gz) return;
    if (gy && !wrap_y) return;
    if (gz && !wrap_z) return;
    const int js = gy ? (j < 1 ? j + L.ny : j - L.ny) : j;
    const int ks = gz ? (k < 1 ? k + L.nz : k - L.nz) : k;
    d.phi[L.at(i, j, k)] = phi_now(d, L.at(i, js, ks));
}

// ------------------------------------------------------------------------------------------------------
// inlet / outlet / porous plate (Boundary_multiphase_inlet.F90, _outlet.F90, _other.F90), one thread per
// (i, j) of the face; `x * (1 - w) + y * w` blends are kept literally (a solid face node keeps its value)
// ------------------------------------------------------------------------------------------------------

__device__ __forceinline__ void set_inlet_phi(const Dev& d, int i, int j, double phi_inlet, int wi) {
    const Layout L = d.L;
    double* ph = d.phi;
    const double v = phi_inlet * (double)(1 - wi) + ph[L.at(i, j, 0)] * (double)wi;
    ph[L.at(i, j, 0)] = v;
    ph[L.at(i, j, -1)] = v;
    ph[L.at(i, j, -2)] = v;
    ph[L.at(i, j, -3)] = v;
}

// inlet_bounce_back_velocity_BC_before_odd (Boundary_multiphase_inlet.F90:6-53)
template <class T, bool HOIST = true>
__device__ __forceinline__ void bc_inlet_velocity_before(const Dev& d, int t, double phi_inlet, double sa_inject, double relaxation) {
    const Layout L = d.L;
    if (t >= L.nx * L.ny) return;
    const int i = t % L.nx + 1, j = t / L.nx + 1;
    // slot written at z = 0 (direction QW, offset DI, DJ) from the slot QR of (i, j, 1); every slot index is
    // resolved first and every slot loaded before the first store, so the latency-bound accesses overlap
    const int QW[5] = {5, 11, 12, 15, 16}, QR[5] = {6, 14, 13, 18, 17}, DI[5] = {0, -1, 1, 0, 0}, DJ[5] = {0, 0, 0, -1, 1};
    int iw[5], ir[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        iw[m] = slot_index(d, QW[m], L.at(i + DI[m], j + DJ[m], 0));
        ir[m] = slot_index(d, QR[m], L.at(i, j, 1));
    }
    FaceSlot<T, HOIST> W[5], R[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        W[m].load(d, QW[m], iw[m]);
        R[m].load(d, QR[m], ir[m]);
    }
    const int wi = d.walls[L.at(i, j, 1)];
    const double om = (double)(1 - wi), wd = (double)wi;
    double tmp2 = d.w_in[plane_at(d, i, j)] * relaxation;
    set_inlet_phi(d, i, j, phi_inlet, wi);
    const double tmp1 = tmp2 * sa_inject;
    tmp2 = tmp2 - tmp1;
    const double c1 = 6.0 * lat::W1, c2 = 6.0 * lat::W2;
    for (int c = 0; c < 2; ++c) {
        const double s = c == 0 ? tmp1 : tmp2;
#pragma unroll
        for (int m = 0; m < 5; ++m) W[m].set(c, QW[m], (R[m].get(c, QR[m]) + (m == 0 ? c1 : c2) * s) * om + W[m].get(c, QW[m]) * wd);
    }
#pragma unroll
    for (int m = 0; m < 5; ++m) W[m].store(d, QW[m], iw[m]);
}
// inlet_bounce_back_velocity_BC_after_odd (Boundary_multiphase_inlet.F90:56-102)
template <class T, bool HOIST = true>
__device__ __forceinline__ void bc_inlet_velocity_after(const Dev& d, int t, double phi_inlet, double sa_inject, double relaxation) {
    const Layout L = d.L;
    if (t >= L.nx * L.ny) return;
    const int i = t % L.nx + 1, j = t / L.nx + 1;
    const int o = L.at(i, j, 1);
    // slot QW of (i, j, 1) from the slot QR at z = 0 (offset DI, DJ); slot indices resolved and slots loaded first
    const int QW[5] = {6, 13, 14, 17, 18}, QR[5] = {5, 12, 11, 16, 15}, DI[5] = {0, 1, -1, 0, 0}, DJ[5] = {0, 0, 0, 1, -1};
    int iw[5], ir[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        iw[m] = slot_index(d, QW[m], o);
        ir[m] = slot_index(d, QR[m], L.at(i + DI[m], j + DJ[m], 0));
    }
    FaceSlot<T, HOIST> W[5], R[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        W[m].load(d, QW[m], iw[m]);
        R[m].load(d, QR[m], ir[m]);
    }
    const int wi = d.walls[L.at(i, j, 1)];
    const double om = (double)(1 - wi), wd = (double)wi;
    double tmp2 = d.w_in[plane_at(d, i, j)] * relaxation;
    set_inlet_phi(d, i, j, phi_inlet, wi);
    const double tmp1 = tmp2 * sa_inject;
    tmp2 = tmp2 - tmp1;
    const double c1 = 6.0 * lat::W1, c2 = 6.0 * lat::W2;
    for (int c = 0; c < 2; ++c) {
        const double s = c == 0 ? tmp1 : tmp2;
#pragma unroll
        for (int m = 0; m < 5; ++m) W[m].set(c, QW[m], (R[m].get(c, QR[m]) + (m == 0 ? c1 : c2) * s) * om + W[m].get(c, QW[m]) * wd);
    }
#pragma unroll
    for (int m = 0; m < 5; ++m) W[m].store(d, QW[m], iw[m]);
}
// inlet_Zou_He_pressure_BC_before_odd (Boundary_multiphase_inlet.F90:115-209)
template <class T>
__device__ __forceinline__ void bc_inlet_pressure_before(const Dev& d, int t, double phi_inlet, double sa_inject, double rho_in, double relaxation) {
    const Layout L = d.L;
    if (t >= L.nx * L.ny) return;
    const int i = t % L.nx + 1, j = t / L.nx + 1;
    const int wi = d.walls[L.at(i, j, 1)];
    const double om = (double)(1 - wi), wd = (double)wi;
    set_inlet_phi(d, i, j, phi_inlet, wi);
    double tmpRho2 = rho_in;
    const double tmpRho1 = rho_in * sa_inject;
    tmpRho2 = tmpRho2 - tmpRho1;
    for (int c = 0; c < 2; ++c) {
        const Pop<T> A = c == 0 ? fpop<T>(d) : gpop<T>(d);
        const double rr = c == 0 ? tmpRho1 : tmpRho2;
        const double s = (rr - (A(0, L.at(i, j, 1)) + A(1, L.at(i - 1, j, 1)) + A(2, L.at(i + 1, j, 1)) +
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
}
// inlet_Zou_He_pressure_BC_after_odd (Boundary_multiphase_inlet.F90:212-294)
template <class T>
__device__ __forceinline__ void bc_inlet_pressure_after(const Dev& d, int t, double phi_inlet, double sa_inject, double rho_in, double relaxation) {
    const Layout L = d.L;
    if (t >= L.nx * L.ny) return;
    const int i = t % L.nx + 1, j = t / L.nx + 1;
    const int wi = d.walls[L.at(i, j, 1)];
    const double om = (double)(1 - wi), wd = (double)wi;
    set_inlet_phi(d, i, j, phi_inlet, wi);
    double tmpRho2 = rho_in;
    const double tmpRho1 = rho_in * sa_inject;
    tmpRho2 = tmpRho2 - tmpRho1;
    const int o = L.at(i, j, 1);
    for (int c = 0; c < 2; ++c) {
        const Pop<T> A = c == 0 ? fpop<T>(d) : gpop<T>(d);
        const double rr = c == 0 ? tmpRho1 : tmpRho2;
        const double s = (rr - (A(0, o) + A(2, o) + A(1, o) + A(4, o) + A(3, o) + A(8, o) + A(7, o) + A(10, o) + A(9, o) +
                                2.0 * (A(5, o) + A(11, o) + A(12, o) + A(15, o) + A(16, o)))) * relaxation;
        const double tnx = 0.5 * (A(2, o) + A(8, o) + A(10, o) - (A(1, o) + A(7, o) + A(9, o)));
        const double tny = 0.5 * (A(4, o) + A(9, o) + A(10, o) - (A(3, o) + A(8, o) + A(7, o)));
        A(6, o) = (A(5, o) + lat::ONE3 * s) * om + A(6, o) * wd;
        A(13, o) = (A(12, o) + lat::ONE6 * s + tnx) * om + A(13, o) * wd;
        A(14, o) = (A(11, o) + lat::ONE6 * s - tnx) * om + A(14, o) * wd;
        A(17, o) = (A(16, o) + lat::ONE6 * s + tny) * om + A(17, o) * wd;
        A(18, o) = (A(15, o) + lat::ONE6 * s - tny) * om + A(18, o) * wd;
    }
}

// outlet_convective_BC_before_odd (Boundary_multiphase_outlet.F90:7-65)
template <class T, bool HOIST = true>
__device__ __forceinline__ void bc_outlet_convective_before(const Dev& d, int t, double u_convec) {
    const Layout L = d.L;
    if (t >= L.nx * L.ny) return;
    const int i = t % L.nx + 1, j = t / L.nx + 1, nz = L.nz;
    const int pc = plane_at(d, i, j);
    const long pl = (long)(L.nx + 2) * (L.ny + 2);
    // (slot, x offset, y offset) of the five unknown populations at the outlet halo; slot indices resolved, and
    // slots and history loaded, before the first store
    const int QS[5] = {6, 13, 14, 17, 18};
    const int DI[5] = {0, -1, 1, 0, 0};
    const int DJ[5] = {0, 0, 0, -1, 1};
    int io[5], ii[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        io[m] = slot_index(d, QS[m], L.at(i + DI[m], j + DJ[m], nz + 1));
        ii[m] = slot_index(d, QS[m], L.at(i + DI[m], j + DJ[m], nz));
    }
    FaceSlot<T, HOIST> O[5], I[5];
    double H[2][5];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        O[m].load(d, QS[m], io[m]);
        I[m].load(d, QS[m], ii[m]);
        if constexpr (HOIST) {
            H[0][m] = d.cf[pc + QS[m] * pl];
            H[1][m] = d.cg[pc + QS[m] * pl];
        }
    }
    const double temp = 1.0 / (1.0 + u_convec);
    const int wi = d.walls[L.at(i, j, nz)];
    const double om = (double)(1 - wi), wd = (double)wi;
    double* ph = d.phi;
    const double v = ((d.cphi[pc] + u_convec * phi_now(d, L.at(i, j, nz))) * temp) * om + ph[L.at(i, j, nz + 1)] * wd;
    d.cphi[pc] = v;
    ph[L.at(i, j, nz + 1)] = v;
    ph[L.at(i, j, nz + 2)] = v;
    ph[L.at(i, j, nz + 3)] = v;
    ph[L.at(i, j, nz + 4)] = v;
    for (int c = 0; c < 2; ++c) {
        double* hist = c == 0 ? d.cf : d.cg;
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            double h;
            if constexpr (HOIST) h = H[c][m];
            else h = hist[pc + QS[m] * pl];
            O[m].set(c, QS[m], ((h + u_convec * I[m].get(c, QS[m])) * temp) * om + O[m].get(c, QS[m]) * wd);
        }
#pragma unroll
        for (int m = 0; m < 5; ++m) hist[pc + QS[m] * pl] = O[m].get(c, QS[m]);
    }
#pragma unroll
    for (int m = 0; m < 5; ++m) O[m].store(d, QS[m], io[m]);
}
// outlet_convective_BC_after_odd (Boundary_multiphase_outlet.F90:68-127)
template <class T, bool HOIST = true>
__device__ __forceinline__ void bc_outlet_convective_after(const Dev& d, int t, double u_convec) {
    const Layout L = d.L;
    if (t >= L.nx * L.ny) return;
    const int i = t % L.nx + 1, j = t / L.nx + 1, nz = L.nz;
    const int pc = plane_at(d, i, j);
    const long pl = (long)(L.nx + 2) * (L.ny + 2);
    // slot written at z = nz and the history slot it relaxes toward; slot indices resolved, and slots and history
    // loaded, before the first store
    const int QW[5] = {5, 11, 12, 15, 16};
    const int QH[5] = {6, 14, 13, 18, 17};
    const int o = L.at(i, j, nz), om1 = L.at(i, j, nz - 1);
    int iw[5], im[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        iw[m] = slot_index(d, QW[m], o);
        im[m] = slot_index(d, QW[m], om1);
    }
    FaceSlot<T, HOIST> W[5], M[5];
    double H[2][5];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        W[m].load(d, QW[m], iw[m]);
        M[m].load(d, QW[m], im[m]);
        if constexpr (HOIST) {
            H[0][m] = d.cf[pc + QH[m] * pl];
            H[1][m] = d.cg[pc + QH[m] * pl];
        }
    }
    const double temp = 1.0 / (1.0 + u_convec);
    const int wi = d.walls[L.at(i, j, nz)];
    const double om = (double)(1 - wi), wd = (double)wi;
    double* ph = d.phi;
    const double v = ((d.cphi[pc] + u_convec * phi_now(d, L.at(i, j, nz))) * temp) * om + ph[L.at(i, j, nz + 1)] * wd;
    d.cphi[pc] = v;
    ph[L.at(i, j, nz + 1)] = v;
    ph[L.at(i, j, nz + 2)] = v;
    ph[L.at(i, j, nz + 3)] = v;
    ph[L.at(i, j, nz + 4)] = v;
    for (int c = 0; c < 2; ++c) {
        double* hist = c == 0 ? d.cf : d.cg;
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            double h;
            if constexpr (HOIST) h = H[c][m];
            else h = hist[pc + QH[m] * pl];
            W[m].set(c, QW[m], ((h + u_convec * M[m].get(c, QW[m])) * temp) * om + W[m].get(c, QW[m]) * wd);
        }
#pragma unroll
        for (int m = 0; m < 5; ++m) hist[pc + QH[m] * pl] = W[m].get(c, QW[m]);
    }
#pragma unroll
    for (int m = 0; m < 5; ++m) W[m].store(d, QW[m], iw[m]);
}
// outlet_Zou_He_pressure_BC_before_odd (Boundary_multiphase_outlet.F90:139-228)
template <class T>
__device__ __forceinline__ void bc_outlet_pressure_before(const Dev& d, int t, double rho_out) {
    const Layout L = d.L;
    if (t >= L.nx * L.ny) return;
    const int i = t % L.nx + 1, j = t / L.nx + 1, nz = L.nz;
    const int wi = d.walls[L.at(i, j, nz)];
    const double om = 1.0 - (double)wi, wd = (double)wi;
    double* ph = d.phi;
    const double pz = phi_now(d, L.at(i, j, nz));
    ph[L.at(i, j, nz + 1)] = pz;
    ph[L.at(i, j, nz + 2)] = pz;
    ph[L.at(i, j, nz + 3)] = pz;
    ph[L.at(i, j, nz + 4)] = pz;
    const Pop<T> f = fpop<T>(d), g = gpop<T>(d);
    const int a0 = L.at(i, j, nz), am = L.at(i - 1, j, nz), ap = L.at(i + 1, j, nz), bm = L.at(i, j - 1, nz),
              bp = L.at(i, j + 1, nz), mm = L.at(i - 1, j - 1, nz), pm = L.at(i + 1, j - 1, nz),
              mp = L.at(i - 1, j + 1, nz), pp = L.at(i + 1, j + 1, nz);
    const int c1 = L.at(i, j, nz - 1), c2 = L.at(i - 1, j, nz - 1), c3 = L.at(i + 1, j, nz - 1), c4 = L.at(i, j - 1, nz - 1),
              c5 = L.at(i, j + 1, nz - 1);
    double tmp1 = (f(0, a0) + f(1, am) + f(2, ap) + f(3, bm) + f(4, bp) + f(7, mm) + f(8, pm) + f(9, mp) + f(10, pp) +
                   2.0 * (f(5, c1) + f(11, c2) + f(12, c3) + f(15, c4) + f(16, c5)) + g(0, a0) + g(1, am) + g(2, ap) +
                   g(3, bm) + g(4, bp) + g(7, mm) + g(8, pm) + g(9, mp) + g(10, pp) +
                   2.0 * (g(5, c1) + g(11, c2) + g(12, c3) + g(15, c4) + g(16, c5))) - rho_out;
    const double tmp2 = tmp1 * 0.5 * (1.0 - pz);
    tmp1 = tmp1 - tmp2;
    for (int c = 0; c < 2; ++c) {
        const Pop<T> A = c == 0 ? f : g;
        const double s = c == 0 ? tmp1 : tmp2;
        const double tnx = 0.5 * (A(1, am) + A(7, mm) + A(9, mp) - (A(2, ap) + A(8, pm) + A(10, pp)));
        const double tny = 0.5 * (A(3, bm) + A(7, mm) + A(8, pm) - (A(4, bp) + A(10, pp) + A(9, mp)));
        int o;
        o = L.at(i, j, nz + 1);     A(6, o) = (A(5, c1) - lat::ONE3 * s) * om + A(6, o) * wd;
        o = L.at(i - 1, j, nz + 1); A(13, o) = (A(12, c3) - lat::ONE6 * s - tnx) * om + A(13, o) * wd;
        o = L.at(i + 1, j, nz + 1); A(14, o) = (A(11, c2) - lat::ONE6 * s + tnx) * om + A(14, o) * wd;
        o = L.at(i, j - 1, nz + 1); A(17, o) = (A(16, c5) - lat::ONE6 * s - tny) * om + A(17, o) * wd;
        o = L.at(i, j + 1, nz + 1); A(18, o) = (A(15, c4) - lat::ONE6 * s + tny) * om + A(18, o) * wd;
    }
}
// outlet_Zou_He_pressure_BC_after_odd (Boundary_multiphase_outlet.F90:231-311)
template <class T>
__device__ __forceinline__ void bc_outlet_pressure_after(const Dev& d, int t, double rho_out) {
    const Layout L = d.L;
    if (t >= L.nx * L.ny) return;
    const int i = t % L.nx + 1, j = t / L.nx + 1, nz = L.nz;
    const int wi = d.walls[L.at(i, j, nz)];
    const double om = 1.0 - (double)wi, wd = (double)wi;
    double* ph = d.phi;
    const double pz = phi_now(d, L.at(i, j, nz));
    ph[L.at(i, j, nz + 1)] = pz;
    ph[L.at(i, j, nz + 2)] = pz;
    ph[L.at(i, j, nz + 3)] = pz;
    ph[L.at(i, j, nz + 4)] = pz;
    const Pop<T> f = fpop<T>(d), g = gpop<T>(d);
    const int o = L.at(i, j, nz);
    double tmp1 = (f(0, o) + f(2, o) + f(1, o) + f(4, o) + f(3, o) + f(8, o) + f(7, o) + f(10, o) + f(9, o) +
                   2.0 * (f(6, o) + f(14, o) + f(13, o) + f(18, o) + f(17, o)) + g(0, o) + g(2, o) + g(1, o) + g(4, o) +
                   g(3, o) + g(8, o) + g(7, o) + g(10, o) + g(9, o) +
                   2.0 * (g(6, o) + g(14, o) + g(13, o) + g(18, o) + g(17, o))) - rho_out;
    const double tmp2 = tmp1 * 0.5 * (1.0 - pz);
    tmp1 = tmp1 - tmp2;
    for (int c = 0; c < 2; ++c) {
        const Pop<T> A = c == 0 ? f : g;
        const double s = c == 0 ? tmp1 : tmp2;
        const double tnx = 0.5 * (A(2, o) + A(8, o) + A(10, o) - (A(1, o) + A(7, o) + A(9, o)));
        const double tny = 0.5 * (A(4, o) + A(10, o) + A(9, o) - (A(3, o) + A(7, o) + A(8, o)));
        A(5, o) = (A(6, o) - lat::ONE3 * s) * om + A(5, o) * wd;
        A(11, o) = (A(14, o) - lat::ONE6 * s + tnx) * om + A(11, o) * wd;
        A(12, o) = (A(13, o) - lat::ONE6 * s - tnx) * om + A(12, o) * wd;
        A(15, o) = (A(18, o) - lat::ONE6 * s + tny) * om + A(15, o) * wd;
        A(16, o) = (A(17, o) - lat::ONE6 * s - tny) * om + A(16, o) * wd;
    }
}

// one launch per boundary condition (the porous-plate and short-domain cases)
template <class T>
__global__ void __launch_bounds__(BLOCK) k_inlet_velocity_before(Dev d, double phi_inlet, double sa_inject, double relaxation) {
    bc_inlet_velocity_before<T>(d, blockIdx.x * BLOCK + threadIdx.x, phi_inlet, sa_inject, relaxation);
}
template <class T>
__global__ void __launch_bounds__(BLOCK) k_inlet_velocity_after(Dev d, double phi_inlet, double sa_inject, double relaxation) {
    bc_inlet_velocity_after<T>(d, blockIdx.x * BLOCK + threadIdx.x, phi_inlet, sa_inject, relaxation);
}
template <class T>
__global__ void __launch_bounds__(BLOCK) k_inlet_pressure_before(Dev d, double phi_inlet, double sa_inject, double rho_in, double relaxation) {
    bc_inlet_pressure_before<T>(d, blockIdx.x * BLOCK + threadIdx.x, phi_inlet, sa_inject, rho_in, relaxation);
}
template <class T>
__global__ void __launch_bounds__(BLOCK) k_inlet_pressure_after(Dev d, double phi_inlet, double sa_inject, double rho_in, double relaxation) {
    bc_inlet_pressure_after<T>(d, blockIdx.x * BLOCK + threadIdx.x, phi_inlet, sa_inject, rho_in, relaxation);
}
template <class T>
__global__ void __launch_bounds__(BLOCK) k_outlet_convective_before(Dev d, double u_convec) {
    bc_outlet_convective_before<T>(d, blockIdx.x * BLOCK + threadIdx.x, u_convec);
}
template <class T>
__global__ void __launch_bounds__(BLOCK) k_outlet_convective_after(Dev d, double u_convec) {
    bc_outlet_convective_after<T>(d, blockIdx.x * BLOCK + threadIdx.x, u_convec);
}
template <class T>
__global__ void __launch_bounds__(BLOCK) k_outlet_pressure_before(Dev d, double rho_out) {
    bc_outlet_pressure_before<T>(d, blockIdx.x * BLOCK + threadIdx.x, rho_out);
}
template <class T>
__global__ void __launch_bounds__(BLOCK) k_outlet_pressure_after(Dev d, double rho_out) {
    bc_outlet_pressure_after<T>(d, blockIdx.x * BLOCK + threadIdx.x, rho_out);
}

// Inlet and outlet conditions of one step in one launch: blocks [0, gface) run the inlet condition, blocks
// [gface, 2 gface) the outlet condition.  They touch disjoint planes (k <= 2 and k >= nz - 1; the launcher
// requires nz >= 8), so running them together equals the reference's inlet-then-outlet order; both are
// latency-bound plane kernels, and one launch overlaps them.  IN / OUT: 1 velocity / convective, 2 pressure.
struct FaceArgs {
    double phi_inlet, sa_inject, relaxation, rho_in, u_convec, rho_out;
};
template <class T, int IN, int OUT, bool AFTER, bool HOIST>
__device__ __forceinline__ void open_faces_block(const Dev& d, const FaceArgs& a, int gface, int b) {
    if (b < gface) {
        const int t = b * BLOCK + (int)threadIdx.x;
        if (IN == 1) {
            if (AFTER) bc_inlet_velocity_after<T, HOIST>(d, t, a.phi_inlet, a.sa_inject, a.relaxation);
            else bc_inlet_velocity_before<T, HOIST>(d, t, a.phi_inlet, a.sa_inject, a.relaxation);
        } else {
            if (AFTER) bc_inlet_pressure_after<T>(d, t, a.phi_inlet, a.sa_inject, a.rho_in, a.relaxation);
            else bc_inlet_pressure_before<T>(d, t, a.phi_inlet, a.sa_inject, a.rho_in, a.relaxation);
        }
    } else {
        const int t = (b - gface) * BLOCK + (int)threadIdx.x;
        if (OUT == 1) {
            if (AFTER) bc_outlet_convective_after<T, HOIST>(d, t, a.u_convec);
            else bc_outlet_convective_before<T, HOIST>(d, t, a.u_convec);
        } else {
            if (AFTER) bc_outlet_pressure_after<T>(d, t, a.rho_out);
            else bc_outlet_pressure_before<T>(d, t, a.rho_out);
        }
    }
}
template <class T, int IN, int OUT, bool AFTER>
__global__ void __launch_bounds__(BLOCK) k_open_faces(Dev d, FaceArgs a, int gface) {
    open_faces_block<T, IN, OUT, AFTER, true>(d, a, gface, (int)blockIdx.x);
}

// porous_plate_BC_before_odd / _after_odd (Boundary_multiphase_other.F90:8-184); `blocked` is the colour
// reflected by the plate (0 = fluid 1, 1 = fluid 2), the other colour passes straight through
template <class T>
__global__ void k_porous_plate(Dev d, int zp, int blocked, int after) {
    const Layout L = d.L;
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= L.nx * L.ny) return;
    const int i = t % L.nx + 1, j = t / L.nx + 1;
    const Pop<T> B = blocked == 0 ? fpop<T>(d) : gpop<T>(d);
    const Pop<T> P = blocked == 0 ? gpop<T>(d) : fpop<T>(d);
    if (!after) {
        B(6, L.at(i, j, zp)) = B(5, L.at(i, j, zp - 1));
        B(13, L.at(i - 1, j, zp)) = B(12, L.at(i, j, zp - 1));
        B(14, L.at(i + 1, j, zp)) = B(11, L.at(i, j, zp - 1));
        B(17, L.at(i, j - 1, zp)) = B(16, L.at(i, j, zp - 1));
        B(18, L.at(i, j + 1, zp)) = B(15, L.at(i, j, zp - 1));
        B(5, L.at(i, j, zp)) = B(6, L.at(i, j, zp + 1));
        B(12, L.at(i + 1, j, zp)) = B(13, L.at(i, j, zp + 1));
        B(11, L.at(i - 1, j, zp)) = B(14, L.at(i, j, zp + 1));
        B(16, L.at(i, j + 1, zp)) = B(17, L.at(i, j, zp + 1));
        B(15, L.at(i, j - 1, zp)) = B(18, L.at(i, j, zp + 1));
        const int QU[5] = {6, 13, 14, 17, 18}, QD[5] = {5, 12, 11, 16, 15};
        for (int m = 0; m < 5; ++m) P(QU[m], L.at(i, j, zp)) = P(QU[m], L.at(i, j, zp + 1));
        for (int m = 0; m < 5; ++m) P(QD[m], L.at(i, j, zp)) = P(QD[m], L.at(i, j, zp - 1));
    } else {
        B(5, L.at(i, j, zp - 1)) = B(6, L.at(i, j, zp));
        B(11, L.at(i, j, zp - 1)) = B(14, L.at(i + 1, j, zp));
        B(12, L.at(i, j, zp - 1)) = B(13, L.at(i - 1, j, zp));
        B(15, L.at(i, j, zp - 1)) = B(18, L.at(i, j + 1, zp));
        B(16, L.at(i, j, zp - 1)) = B(17, L.at(i, j - 1, zp));
        B(6, L.at(i, j, zp + 1)) = B(5, L.at(i, j, zp));
        B(14, L.at(i, j, zp + 1)) = B(11, L.at(i - 1, j, zp));
        B(13, L.at(i, j, zp + 1)) = B(12, L.at(i + 1, j, zp));
        B(18, L.at(i, j, zp + 1)) = B(15, L.at(i, j - 1, zp));
        B(17, L.at(i, j, zp + 1)) = B(16, L.at(i, j + 1, zp));
        const int QD[5] = {5, 11, 12, 15, 16}, QU[5] = {6, 14, 13, 18, 17};
        for (int m = 0; m < 5; ++m) P(QD[m], L.at(i, j, zp - 1)) = P(QD[m], L.at(i, j, zp));
        for (int m = 0; m < 5; ++m) P(QU[m], L.at(i, j, zp + 1)) = P(QU[m], L.at(i, j, zp));
    }
}

// ------------------------------------------------------------------------------------------------------
// colour-gradient chain (color_gradient, Phase_gradient.F90:5-204), each step over a node list
// ------------------------------------------------------------------------------------------------------

// solid boundary nodes: padded offset, bit q set when neighbour q is a fluid neighbour of the reference's
// neighbour_list (which lists them in increasing q), and the extrapolation weight la_weight
struct SolidList {
    int n;
    const int* pos;
    const unsigned* mask;
    const double* w;
};
// fluid boundary nodes: padded offset, wall normal, cos/sin of the contact angle
struct FluidList {
    int n;
    const int* pos;
    const double* nw;    // 3 per node
    const double* cos_t;
    const double* sin_t;
};
struct NodeList {
    int n;
    const int* pos;
};

// ISO4 first derivatives with the reference's term order (Phase_gradient.F90:40-62).  A(dx, dy, dz) is the
// field at the offset (as a double); the same expressions serve the dense arrays and the LDS tiles of the fused
// chain, so both evaluate identical operations in identical order.
template <class F>
__device__ __forceinline__ double dx4f(const F& A) {
    return lat::ISO4_1 * (A(1, 0, 0) - A(-1, 0, 0)) +
           lat::ISO4_2 * (A(1, 1, 0) - A(-1, -1, 0) + A(1, -1, 0) - A(-1, 1, 0) + A(1, 0, 1) - A(-1, 0, -1) +
                          A(1, 0, -1) - A(-1, 0, 1));
}
template <class F>
__device__ __forceinline__ double dy4f(const F& A) {
    return lat::ISO4_1 * (A(0, 1, 0) - A(0, -1, 0)) +
           lat::ISO4_2 * (A(1, 1, 0) - A(-1, -1, 0) + A(-1, 1, 0) - A(1, -1, 0) + A(0, 1, 1) - A(0, -1, -1) +
                          A(0, 1, -1) - A(0, -1, 1));
}
template <class F>
__device__ __forceinline__ double dz4f(const F& A) {
    return lat::ISO4_1 * (A(0, 0, 1) - A(0, 0, -1)) +
           lat::ISO4_2 * (A(1, 0, 1) - A(-1, 0, -1) + A(-1, 0, 1) - A(1, 0, -1) + A(0, 1, 1) - A(0, -1, -1) +
                          A(0, -1, 1) - A(0, 1, -1));
}
template <class T>
struct DenseAt {
    const T* A;
    int n, px, pxy;
    __device__ __forceinline__ double operator()(int a, int b, int c) const { return (double)A[n + a + b * px + c * pxy]; }
};
template <class T>
__device__ __forceinline__ double dx4(const T* A, int n, int px, int pxy) { return dx4f(DenseAt<T>{A, n, px, pxy}); }
template <class T>
__device__ __forceinline__ double dy4(const T* A, int n, int px, int pxy) { return dy4f(DenseAt<T>{A, n, px, pxy}); }
template <class T>
__device__ __forceinline__ double dz4(const T* A, int n, int px, int pxy) { return dz4f(DenseAt<T>{A, n, px, pxy}); }
// curvature from the normal components X, Y, Z as accessors (Phase_gradient.F90:112-201)
template <class FX, class FY, class FZ>
__device__ __forceinline__ double curvature_f(const FX& X, const FY& Y, const FZ& Z) {
    const double kxx = dx4f(X);
    const double kyy = dy4f(Y);
    const double kzz = dz4f(Z);
    const double kxy = dy4f(X);
    const double kxz = dz4f(X);
    const double kyx = dx4f(Y);
    const double kyz = dz4f(Y);
    const double kzx = dx4f(Z);
    const double kzy = dy4f(Z);
    const double cx = X(0, 0, 0), cy = Y(0, 0, 0), cz = Z(0, 0, 0);
    return (cx * cx - 1.0) * kxx + (cy * cy - 1.0) * kyy + (cz * cz - 1.0) * kzz + cx * cy * (kxy + kyx) +
           cx * cz * (kxz + kzx) + cy * cz * (kzy + kyz);
}
// interface normal and |grad phi| from the gradient (Phase_gradient.F90:33-79): below 1e-6 everything is 0
__device__ __forceinline__ void normalise(double& cx, double& cy, double& cz, double& cn) {
    cn = sqrt(cx * cx + cy * cy + cz * cz);
    if (cn < 1e-6) {
        cx = 0.0;
        cy = 0.0;
        cz = 0.0;
        cn = 0.0;
    } else {
        cx = cx / cn;
        cy = cy / cn;
        cz = cz / cn;
    }
}

// The normal field: FP64 (the reference arithmetic) or, in the FP32 storage modes, FP32 (the gradient and
// curvature are still evaluated in FP64 from the stored values; the FP32 field halves the bytes the chain and
// the lattice update's curvature stencil move).  Rec<NT>: the record type.
template <class NT>
struct Rec;
template <>
struct Rec<double> {
    using type = double4;
};
template <>
struct Rec<float> {
    using type = float4;
};
template <class NT>
using Rec4 = typename Rec<NT>::type;
template <class NT>
__device__ __forceinline__ Rec4<NT>* normals(const Dev& d);
template <>
__device__ __forceinline__ double4* normals<double>(const Dev& d) { return d.cn; }
template <>
__device__ __forceinline__ float4* normals<float>(const Dev& d) { return d.cn32; }
template <class NT>
__device__ __forceinline__ Rec4<NT> rec4(double x, double y, double z, double n) {
    return Rec4<NT>{(NT)x, (NT)y, (NT)z, (NT)n};
}

// any bit of the stored value set (-0.0 and NaN count as set)
__device__ __forceinline__ bool bits_set(double v) { return __double_as_longlong(v) != 0; }
__device__ __forceinline__ bool bits_set(float v) { return __float_as_uint(v) != 0u; }
struct ZeroAt {
    __device__ __forceinline__ double operator()(int, int, int) const { return 0.0; }
};
__device__ __forceinline__ double quiet_curvature() { return curvature_f(ZeroAt{}, ZeroAt{}, ZeroAt{}); }

// phi extrapolated into a solid boundary node from its fluid neighbours (Phase_gradient.F90:12-30).  Solid list
// nodes lie within [-2, n+3]^3, so all 18 neighbours are inside the padded grid.  The 18 loads are issued together,
// without branches (phi_now: the current phi also where the dense phi lags); a direction without a fluid neighbour
// loads the node's own value instead (a line already fetched, the value unused), and the fluid ones are summed in
// the reference's order.
template <bool DENSE = false>
__device__ __forceinline__ void phi_extrapolate_at(const Dev& d, const SolidList& s, int t) {
    const int n = s.pos[t];
    const unsigned mk = s.mask[t];
    double v[19];
#pragma unroll
    for (int q = 1; q < 19; ++q) v[q] = phi_at<DENSE>(d, (mk & (1u << q)) ? n + d.L.off(q) : n);
    double acc = 0.0;
#pragma unroll
    for (int q = 1; q < 19; ++q) acc = (mk & (1u << q)) ? acc + v[q] * lat::W[q] : acc;
    d.phi[n] = acc / s.w[t];
}
// every entry of the list, from the dense phi (made current first): before a download (the brick chain skips the
// bricks no evaluated normal reads) and in the whole-list chain
__global__ void __launch_bounds__(BLOCK) k_phi_extrapolate(Dev d, SolidList s, int per) {
    const int t = xcd_block(per) * BLOCK + threadIdx.x;
    if (t >= s.n) return;
    phi_extrapolate_at<true>(d, s, t);
}
// wetting: turn the normal at a fluid node next to the wall to the contact angle (Akai et al. 2018;
// alter_color_gradient_solid_surface, Phase_gradient.F90:210-265).  cos/sin of the angle come from the host.
__device__ __forceinline__ void wet(double& cx, double& cy, double& cz, double nwx, double nwy, double nwz, double tmpCos,
                                    double sinT) {
    const double cnx0 = cx, cny0 = cy, cnz0 = cz;
    const double tmp1 = nwx * cnx0 + nwy * cny0 + nwz * cnz0;
    const double tmp2 = 1.0 / sqrt(1.0 - tmp1 * tmp1);
    const double coe1 = sinT * tmp1 * tmp2;
    const double coe2 = sinT * tmp2;
    const double cnxp = (tmpCos - coe1) * nwx + coe2 * cnx0;
    const double cnyp = (tmpCos - coe1) * nwy + coe2 * cny0;
    const double cnzp = (tmpCos - coe1) * nwz + coe2 * cnz0;
    const double cnxm = (tmpCos + coe1) * nwx - coe2 * cnx0;
    const double cnym = (tmpCos + coe1) * nwy - coe2 * cny0;
    const double cnzm = (tmpCos + coe1) * nwz - coe2 * cnz0;
    const double distP = (cnxp - cnx0) * (cnxp - cnx0) + (cnyp - cny0) * (cnyp - cny0) + (cnzp - cnz0) * (cnzp - cnz0);
    const double distM = (cnxm - cnx0) * (cnxm - cnx0) + (cnym - cny0) * (cnym - cny0) + (cnzm - cnz0) * (cnzm - cnz0);
    if (distP <= distM) {
        cx = cnxp;
        cy = cnyp;
        cz = cnzp;
    } else {
        cx = cnxm;
        cy = cnym;
        cz = cnzm;
    }
}
// The chain stages over the bricks a plan (k_brick_plan) listed.  Each list (normal nodes, solid nodes, solid
// nodes within one halo layer) is sorted by brick, off[b] .. off[b + 1] being brick b's entries (at most 64, the
// nodes of a brick).  A group of G lanes takes one listed brick at a time; each group owns a contiguous run of the
// work list (groups in the XCD-aware block order, so a die works on neighbouring bricks), and the grid is fixed:
// the list length is known on the device only.
struct BrickWork {
    const int4* list;   // bricks to visit: (brick, first entry, end entry, brick quiet), NXCD sub-lists of cap items
    const int* cnt;     // their numbers (NXCD)
    int cap;
};
constexpr int CHAIN_BLOCKS = 2048;   // grid of the brick-list kernels: what the GPU holds at once (1024-8192 measured alike)
constexpr int CHAIN_PER = CHAIN_BLOCKS / NXCD;     // work-groups per XCD
static_assert(CHAIN_BLOCKS % NXCD == 0, "whole work-groups per XCD");
// The work items of a group of G lanes.  A work list is kept as NXCD sub-lists (the plan appends a tile's bricks
// to sub-list tile % NXCD, so no counter takes more than its share of the appends); the groups of XCD x take the
// items of sub-list x in turn (work-groups are dispatched to the dies round-robin, blockIdx.x % NXCD).
template <int G>
struct GroupItems {
    static constexpr int NG = BLOCK / G;
    int i, hi, lane;
    const int4* list;
    __device__ __forceinline__ GroupItems(const BrickWork& w, int blk = (int)blockIdx.x) {
        const int x = blk % NXCD;
        hi = w.cnt[x];
        list = w.list + (size_t)x * w.cap;
        i = (blk / NXCD) * NG + (int)threadIdx.x / G;
        lane = (int)threadIdx.x % G;
    }
    static constexpr int STRIDE = CHAIN_PER * NG;
    // the item at i, with the load of the group's next item started (it overlaps this item's loads)
    int4 pre;
    __device__ __forceinline__ void start() {
        if (i < hi) pre = list[i];
    }
    __device__ __forceinline__ int4 take() {
        const int4 it = pre;
        if (i + STRIDE < hi) pre = list[i + STRIDE];
        return it;
    }
};
// a normal of brick b has a bit set when any lane of the group says so: nzb of b and nzd of its 27-brick
// neighbourhood (which the lattice update reads) are set.  All lanes of the wave call this together.
template <int G>
__device__ __forceinline__ void mark_nonzero(const Dev& d, int b, bool nz, int lane) {
    const unsigned long long m = __ballot(nz);
    const unsigned long long gm = G == 64 ? m : (m >> ((threadIdx.x & 63u) & ~(unsigned)(G - 1))) & ((1ull << (G & 63)) - 1ull);
    if (!gm) return;
    if (lane == 0) d.nzb[b] = 1;
    const int bz = b / (d.nbx * d.nby), by = (b / d.nbx) % d.nby, bx = b % d.nbx;
    for (int u = lane; u < 27; u += G) {
        const int x = bx + u % 3 - 1, y = by + (u / 3) % 3 - 1, z = bz + u / 9 - 1;
        if (x >= 0 && x < d.nbx && y >= 0 && y < d.nby && z >= 0 && z < d.nbz) d.nzd[x + d.nbx * (y + d.nby * z)] = 1;
    }
}
// lanes per brick of the two extrapolation stages (a brick has up to 64 solid entries; 16 and 64 lanes measured slower)
constexpr int SOLID_GROUP = 32;
// phi extrapolation at the solid nodes of the listed bricks (those whose phi an evaluated normal may read)
__global__ void __launch_bounds__(BLOCK) k_phi_extrapolate_b(Dev d, SolidList s, BrickWork w) {
    GroupItems<SOLID_GROUP> g(w);
    for (g.start(); g.i < g.hi; g.i += GroupItems<SOLID_GROUP>::STRIDE) {
        const int4 it = g.take();
        for (int t = it.y + g.lane; t < it.z; t += SOLID_GROUP) phi_extrapolate_at(d, s, t);
    }
}
// interface normal and |grad phi| at the non-solid nodes of [-1, n+2]^3 (Phase_gradient.F90:33-79), with the
// wetting rotation applied at once where the node is a fluid boundary node (the rotation reads and writes only
// the node's own normal, so doing it here equals the reference's separate pass).  wb[t] indexes the fluid
// boundary list or is -1.  Solid nodes keep 0 or their extrapolated value, as after the reference's
// zero-then-extrapolate sequence.  Listed bricks are evaluated, or, when q1b (every phi their normals depend on
// lies within the band: the normals are +0), cleared.
// the normal of node n (entry t of the normal list): stored, and whether a bit of it is set
template <class NT, bool DENSE = false>
__device__ __forceinline__ bool eval_normal(const Dev& d, int n, int fb, const FluidList& fl) {
    const PhiNowAt<DENSE> P{&d, n};
    double cx = dx4f(P);
    double cy = dy4f(P);
    double cz = dz4f(P);
    double cn;
    normalise(cx, cy, cz, cn);
    if (fb >= 0 && cn > 1e-6)
        wet(cx, cy, cz, fl.nw[3 * fb], fl.nw[3 * fb + 1], fl.nw[3 * fb + 2], fl.cos_t[fb], fl.sin_t[fb]);
    const NT sx = (NT)cx, sy = (NT)cy, sz = (NT)cz, sn = (NT)cn;
    normals<NT>(d)[n] = Rec4<NT>{sx, sy, sz, sn};
    return bits_set(sx) | bits_set(sy) | bits_set(sz) | bits_set(sn);
}
// the extrapolated normal of solid list entry t: the neighbours loaded nine directions at a time, without branches
// (inside the padded grid, as in phi_extrapolate_at; a direction without a fluid neighbour loads the node's own
// record), and the fluid ones summed in the reference's order; the record is stored whole (c_norm of a solid node is
// never evaluated and stays +0)
template <class NT>
__device__ __forceinline__ bool eval_normal_extrapolate(const Dev& d, const SolidList& s, int t) {
    Rec4<NT>* const N = normals<NT>(d);
    const int n = s.pos[t];
    const unsigned mk = s.mask[t];
    double ax = 0.0, ay = 0.0, az = 0.0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        NT vx[9], vy[9], vz[9];
#pragma unroll
        for (int u = 0; u < 9; ++u) {
            const int q = 1 + 9 * h + u;
            const Rec4<NT> r = N[(mk & (1u << q)) ? n + d.L.off(q) : n];
            vx[u] = r.x;
            vy[u] = r.y;
            vz[u] = r.z;
        }
#pragma unroll
        for (int u = 0; u < 9; ++u) {
            const int q = 1 + 9 * h + u;
            const bool f = (mk & (1u << q)) != 0;
            ax = f ? ax + (double)vx[u] * lat::W[q] : ax;
            ay = f ? ay + (double)vy[u] * lat::W[q] : ay;
            az = f ? az + (double)vz[u] * lat::W[q] : az;
        }
    }
    const NT sx = (NT)(ax / s.w[t]), sy = (NT)(ay / s.w[t]), sz = (NT)(az / s.w[t]);
    N[n] = Rec4<NT>{sx, sy, sz, (NT)0.0};
    return bits_set(sx) || bits_set(sy) || bits_set(sz);
}
template <class NT>
__global__ void __launch_bounds__(BLOCK) k_normals(Dev d, NodeList nl, const int* __restrict__ wb, FluidList fl, BrickWork w) {
    GroupItems<64> g(w);
    Rec4<NT>* const N = normals<NT>(d);
    for (g.start(); g.i < g.hi; g.i += GroupItems<64>::STRIDE) {
        const int4 it = g.take();
        const int t = it.y + g.lane;
        bool nz = false;
        if (t < it.z) {
            const int n = nl.pos[t];
            if (it.w) N[n] = rec4<NT>(0.0, 0.0, 0.0, 0.0);
            else nz = eval_normal<NT>(d, n, wb[t], fl);
        }
        mark_nonzero<64>(d, it.x, nz, g.lane);
    }
}
// normal extrapolated into the solid boundary nodes within one halo layer (Phase_gradient.F90:84-110); listed
// bricks are evaluated, or, when q2b (every normal read is +0, so is the weighted mean), cleared
template <class NT>
__global__ void __launch_bounds__(BLOCK) k_normal_extrapolate(Dev d, SolidList s, BrickWork w) {
    GroupItems<SOLID_GROUP> g(w);
    Rec4<NT>* const N = normals<NT>(d);
    for (g.start(); g.i < g.hi; g.i += GroupItems<SOLID_GROUP>::STRIDE) {
        const int4 it = g.take();
        bool nz = false;
        for (int t = it.y + g.lane; t < it.z; t += SOLID_GROUP) {
            if (it.w) N[s.pos[t]] = rec4<NT>(0.0, 0.0, 0.0, 0.0);
            else nz = eval_normal_extrapolate<NT>(d, s, t) || nz;
        }
        mark_nonzero<SOLID_GROUP>(d, it.x, nz, g.lane);
    }
}
// The whole chain without brick lists (FP32 collision, whose rounding keeps phi out of the band almost everywhere):
// one thread per entry of the normal list and of the solid list within one halo layer.  The lists are in brick
// order, so a wave holds a few runs of entries of one brick each: the first lane of a run with a set bit sets nzd
// of the brick's 27-brick neighbourhood (cleared before the chain).  Every lane of the wave calls this; lanes past
// the end of the list pass b = -1 and nz = false.
__device__ __forceinline__ void mark_nonzero_runs(const Dev& d, int b, bool nz) {
    const unsigned long long m = __ballot(nz);
    if (!m) return;
    const int lane = (int)(threadIdx.x & 63u);
    const int bp = __shfl(b, lane == 0 ? 0 : lane - 1);
    const bool lead = lane == 0 || bp != b;
    const unsigned long long ld = __ballot(lead);
    const unsigned long long above = lane == 63 ? 0ull : ld & (~0ull << (lane + 1));
    const unsigned long long run = above ? (above & (~above + 1ull)) - (1ull << lane) : ~0ull << lane;
    if (!lead || !(m & run)) return;
    const int bz = b / (d.nbx * d.nby), by = (b / d.nbx) % d.nby, bx = b % d.nbx;
    for (int u = 0; u < 27; ++u) {
        const int x = bx + u % 3 - 1, y = by + (u / 3) % 3 - 1, z = bz + u / 9 - 1;
        if (x >= 0 && x < d.nbx && y >= 0 && y < d.nby && z >= 0 && z < d.nbz) d.nzd[x + d.nbx * (y + d.nby * z)] = 1;
    }
}
__global__ void __launch_bounds__(BLOCK) k_clear_bytes(unsigned char* p, int n) {
    const int t = (int)blockIdx.x * BLOCK + (int)threadIdx.x;
    if (t < n) p[t] = 0;
}
template <class NT>
__global__ void __launch_bounds__(BLOCK) k_normals_all(Dev d, NodeList nl, const int* __restrict__ wb, FluidList fl) {
    const int t = (int)blockIdx.x * BLOCK + (int)threadIdx.x;
    int b = -1;
    bool nz = false;
    if (t < nl.n) {
        const int n = nl.pos[t];
        b = brick_of(d, n);
        nz = eval_normal<NT, true>(d, n, wb[t], fl);
    }
    mark_nonzero_runs(d, b, nz);
}
template <class NT>
__global__ void __launch_bounds__(BLOCK) k_normal_extrapolate_all(Dev d, SolidList s) {
    const int t = (int)blockIdx.x * BLOCK + (int)threadIdx.x;
    int b = -1;
    bool nz = false;
    if (t < s.n) {
        b = brick_of(d, s.pos[t]);
        nz = eval_normal_extrapolate<NT>(d, s, t);
    }
    mark_nonzero_runs(d, b, nz);
}
// The plan of one chain.  A work-group owns TBX x TBY x TBZ bricks and stages the band flags of them and a halo
// of two bricks in LDS.  From the band flags the lattice update left it derives q1b (every brick within one quiet
// with one sign; no brick when `full`) over the tile and a one-brick halo, then q2b (q1b over the 27-brick
// neighbourhood).  nzb and nzd are cleared (the chain stages set them again where they store a normal with a bit
// set) and the next band buffers reset to bstatic.  The bricks each stage visits are appended to its work list with
// their entry range: normals where the brick has normal nodes and (!q1b or zfill), normal extrapolation where it
// has solid nodes within one halo layer and (!q2b or zfill), phi extrapolation where it has solid nodes and !q2b;
// zfill is nzb before the clearing: a skipped stage writes +0 where a stored normal has a bit set.  The other set
// of list counters is cleared for the next chain.
// IN != 0: the first 2 gface work-groups run the inlet and outlet conditions of the step instead (k_open_faces):
// they touch no brick flag and read interior phi through phi_now, and the phi they write in the halo planes is
// read only by the chain stages after the plan, so the latency-bound face work overlaps the plan.  ntx, nty: tiles
// per axis.
constexpr int TBX = 8, TBY = 8, TBZ = 4;                     // TBX * TBY * TBZ == BLOCK
constexpr int TSX = TBX + 2, TSY = TBY + 2, TSZ = TBZ + 2;   // q1b region: the tile and a one-brick halo
constexpr int TUX = TBX + 4, TUY = TBY + 4, TUZ = TBZ + 4;   // band region: the tile and a two-brick halo
static_assert(TBX * TBY * TBZ == BLOCK, "one brick per thread");
constexpr int NWORK = 3;   // work lists: normals, normal extrapolation, phi extrapolation
struct ChainPlan {
    const unsigned char* notm;   // band flags of the lattice update since the last chain
    const unsigned char* notp;
    unsigned char* next_m;       // band flags the next lattice update fills
    unsigned char* next_p;
    const unsigned char* hasb;   // per brick: bit 0 normal nodes, bit 1 solid nodes, bit 2 solid nodes in one halo layer
    const int* off_n;            // brick offsets of the sorted lists (nbricks + 1 each)
    const int* off_s;
    const int* off_1;
    int4* list[NWORK];           // work lists (NXCD sub-lists of cap items each) and their counters (NXCD each)
    int cap;
    int* cnt;
    int* cnt_next;               // the counters of the next chain
};
template <class T, int IN, int OUT, bool AFTER>
__global__ void __launch_bounds__(BLOCK) k_brick_plan(Dev d, ChainPlan p, int full, FaceArgs a, int gface, int ntx, int nty) {
    __shared__ unsigned char band[TUX * TUY * TUZ];
    __shared__ unsigned char q1s[TSX * TSY * TSZ];
    __shared__ int wcount[NWORK][BLOCK / 64];
    __shared__ int wbase[NWORK];
    int tile = (int)blockIdx.x;
    if (IN != 0) {
        if (tile < 2 * gface) {
            open_faces_block<T, IN, OUT, AFTER, false>(d, a, gface, tile);
            return;
        }
        tile -= 2 * gface;
    }
    const int bx0 = (tile % ntx) * TBX, by0 = ((tile / ntx) % nty) * TBY, bz0 = (tile / (ntx * nty)) * TBZ;
    auto inside = [&](int x, int y, int z) { return x >= 0 && x < d.nbx && y >= 0 && y < d.nby && z >= 0 && z < d.nbz; };
    const int tx = threadIdx.x % TBX, ty = (threadIdx.x / TBX) % TBY, tz = threadIdx.x / (TBX * TBY);
    const int x = bx0 + tx, y = by0 + ty, z = bz0 + tz;
    const bool valid = x < d.nbx && y < d.nby && z < d.nbz;
    const int b = valid ? x + d.nbx * (y + d.nby * z) : 0;
    // this brick's own flags, loaded with the staging
    const unsigned char zf = valid ? d.nzb[b] : 0, h = valid ? p.hasb[b] : 0, st = valid ? d.bstatic[b] : 0;
    {   // every load of the staging issued before the first is used
        constexpr int NSTAGE = (TUX * TUY * TUZ + BLOCK - 1) / BLOCK;
        unsigned char vm[NSTAGE], vp[NSTAGE];
#pragma unroll
        for (int k = 0; k < NSTAGE; ++k) {
            const int t = threadIdx.x + k * BLOCK;
            const int xx = bx0 + t % TUX - 2, yy = by0 + (t / TUX) % TUY - 2, zz = bz0 + t / (TUX * TUY) - 2;
            const bool in = t < TUX * TUY * TUZ && inside(xx, yy, zz);
            const int c = in ? xx + d.nbx * (yy + d.nby * zz) : 0;
            vm[k] = in ? p.notm[c] : 0;
            vp[k] = in ? p.notp[c] : 0;
        }
#pragma unroll
        for (int k = 0; k < NSTAGE; ++k) {
            const int t = threadIdx.x + k * BLOCK;
            if (t < TUX * TUY * TUZ) band[t] = (unsigned char)((vm[k] ? 1 : 0) | (vp[k] ? 2 : 0));
        }
    }
    __syncthreads();
    // q1b of the region brick (lx - 1, ly - 1, lz - 1) of the tile: its neighbourhood is band (lx..lx+2, ...)
    for (int t = threadIdx.x; t < TSX * TSY * TSZ; t += BLOCK) {
        const int lx = t % TSX, ly = (t / TSX) % TSY, lz = t / (TSX * TSY);
        unsigned char q = 1;   // outside the grid: quiet
        if (inside(bx0 + lx - 1, by0 + ly - 1, bz0 + lz - 1)) {
            unsigned char o = 0;
#pragma unroll
            for (int dz = 0; dz < 3; ++dz)
#pragma unroll
                for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                    for (int dx = 0; dx < 3; ++dx) o |= band[(lx + dx) + TUX * ((ly + dy) + TUY * (lz + dz))];
            const bool m = full != 0 || (o & 1), pl = full != 0 || (o & 2);
            q = (!m || !pl) ? 1 : 0;
        }
        q1s[t] = q;
    }
    __syncthreads();
    bool take[NWORK] = {false, false, false};
    unsigned char q1 = 0, q2 = 0;
    if (valid) {
        const int r = tx + TSX * (ty + TSY * tz);   // this brick's neighbourhood corner in the q1b region
        q1 = q1s[r + 1 + TSX * (1 + TSY)];
        q2 = 1;
#pragma unroll
        for (int dz = 0; dz < 3; ++dz)
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int dx = 0; dx < 3; ++dx) q2 &= q1s[r + dx + TSX * (dy + TSY * dz)];
        d.nzb[b] = 0;
        d.nzd[b] = 0;
        p.next_m[b] = st;   // the next lattice update adds its band flags to the static ones
        p.next_p[b] = st;
        take[0] = (h & 1) && (!q1 || zf);
        take[1] = (h & 4) && (!q2 || zf);
        take[2] = (h & 2) && !q2;
    }
    // one atomic per list and work-group (on the tile's sub-list): wave counts, the group's base, each brick's slot
    const int lane = (int)(threadIdx.x & 63u), wv = (int)threadIdx.x >> 6;
    unsigned long long m[NWORK];
#pragma unroll
    for (int l = 0; l < NWORK; ++l) {
        m[l] = __ballot(take[l]);
        if (lane == 0) wcount[l][wv] = __popcll(m[l]);
    }
    __syncthreads();
    const int sub = tile % NXCD;
    if (threadIdx.x < NWORK) {
        int tot = 0;
        for (int v = 0; v < BLOCK / 64; ++v) tot += wcount[threadIdx.x][v];
        wbase[threadIdx.x] = tot ? atomicAdd(p.cnt + threadIdx.x * NXCD + sub, tot) : 0;
    }
    __syncthreads();
    const int* const offs[NWORK] = {p.off_n, p.off_1, p.off_s};
    const unsigned char quiet[NWORK] = {q1, q2, 0};
#pragma unroll
    for (int l = 0; l < NWORK; ++l) {
        if (!take[l]) continue;
        int slot = wbase[l] + __popcll(m[l] & ((1ull << lane) - 1ull));
        for (int v = 0; v < wv; ++v) slot += wcount[l][v];
        p.list[l][(size_t)sub * p.cap + slot] = int4{b, offs[l][b], offs[l][b + 1], (int)quiet[l]};
    }
    if (tile == 0 && threadIdx.x < NWORK * NXCD) p.cnt_next[threadIdx.x] = 0;
}
// the whole dense phi at the fluid nodes from the compact phi (before a reader outside the chain)
__global__ void k_phi_dense(Dev d) {
    const int m = blockIdx.x * BLOCK + threadIdx.x;
    if (m < d.nf) d.phi[d.cpos[m]] = d.phic[m];
}
// the same, with the n bytes at z cleared in the same launch (the whole-list chain's stencil flags)
__global__ void k_phi_dense_clear(Dev d, unsigned char* z, int n) {
    const int m = blockIdx.x * BLOCK + threadIdx.x;
    if (m < d.nf) d.phi[d.cpos[m]] = d.phic[m];
    if (m < n) z[m] = 0;
}
// the compact phi from the dense phi (after an upload or a device initialisation)
__global__ void k_phic_gather(Dev d) {
    const int m = blockIdx.x * BLOCK + threadIdx.x;
    if (m < d.nf) d.phic[m] = d.phi[d.cpos[m]];
}
// component C (0 x, 1 y, 2 z) of the normal records around node n
template <class NT, int C>
struct RecAt {
    const Rec4<NT>* r;
    int n, px, pxy;
    __device__ __forceinline__ double operator()(int a, int b, int c) const {
        const Rec4<NT>& v = r[n + a + b * px + c * pxy];
        return (double)(C == 0 ? v.x : C == 1 ? v.y : v.z);
    }
};
template <class NT>
__device__ __forceinline__ double curvature_at(const Dev& d, int n) {
    const Rec4<NT>* N = normals<NT>(d);
    const int px = d.L.px, pxy = d.L.pxy;
    return curvature_f(RecAt<NT, 0>{N, n, px, pxy}, RecAt<NT, 1>{N, n, px, pxy}, RecAt<NT, 2>{N, n, px, pxy});
}
// curvature at the fluid nodes (Phase_gradient.F90:112-201), emitted with the normal as the per-node force
// input of the next lattice update: (cn_x, cn_y, cn_z, 0.5*gamma*curv*c_norm), Kernel_multiphase.F90:112-116
template <class NT>
__global__ void __launch_bounds__(BLOCK) k_curvature(Dev d, double gamma, int per) {
    const int m = xcd_block(per) * BLOCK + threadIdx.x;
    if (m >= d.nf) return;
    const int n = d.cpos[m];
    const double curv = curvature_at<NT>(d, n);
    const Rec4<NT> r = normals<NT>(d)[n];
    d.force[m] = (double)r.x;
    d.force[d.fs + m] = (double)r.y;
    d.force[2 * d.fs + m] = (double)r.z;
    d.force[3 * d.fs + m] = 0.5 * gamma * curv * (double)r.w;
}
// The tail of the colour-gradient chain in one pass, for the lattice update's force input: interface normals
// with the wetting rotation, their extrapolation into the solid boundary nodes and the curvature force
// (cn_x, cn_y, cn_z, 0.5 gamma curv c_norm) of every interior fluid node, written to the compact force arrays.
// A work-group owns a CT x CT column of interior nodes and a segment of CSEG planes, and marches along z with
// rolling LDS planes: at step i it loads phi plane i (tile + halo 3), computes the normals of plane i - 1
// (halo 2), extrapolates plane i - 2 (halo 1) and evaluates the curvature of plane i - 3.  Every value is the
// one k_normals, k_normal_extrapolate and k_curvature compute from the same phi (same expressions, same order),
// so the force input is bit-identical to the list kernels'; the dense normal field is not written.
// role[o]: >= 0 fluid boundary index (wetting), <= -2 solid boundary node of the extrapolation list -role - 2.
constexpr int CT = 16;           // output columns per tile edge
constexpr int CSEG = 32;         // planes per work-group
constexpr int CPW = CT + 6;      // phi tile edge (halo 3)
constexpr int CNW = CT + 4;      // normal tile edge (halo 2)
constexpr int CXW = CT + 2;      // extrapolation tile edge (halo 1)
constexpr int CBLOCK = 512;
__device__ __forceinline__ int mod_pos(int a, int m) { return ((a % m) + m) % m; }
template <class NT>
__global__ void __launch_bounds__(CBLOCK) k_chain_force(Dev d, FluidList fl, SolidList s1, const int* __restrict__ role,
                                                        double gamma, int ntx, int nty, int per) {
    __shared__ double sphi[3][CPW * CPW];
    __shared__ NT snrm[4][3][CNW * CNW];   // [plane slot][x, y, z][node]
    __shared__ NT scn[4][CT * CT];         // |grad phi| of the output columns
    __shared__ int8_t sbw[CNW * CNW];      // normals stage: wall flag (1: no normal) of plane i - 1
    __shared__ int sbr[CNW * CNW];         //                role of plane i - 1
    __shared__ int scr[CXW * CXW];         // extrapolation stage: role of plane i - 2 (-1 outside [0, n+1])
    __shared__ int sdm[CT * CT];           // curvature stage: compact index of plane i - 3 (-1: no output)
    const Layout L = d.L;
    const int b = xcd_block(per);
    if (b >= ntx * nty * ((L.nz + CSEG - 1) / CSEG)) return;
    const int tx = b % ntx, ty = (b / ntx) % nty, sg = b / (ntx * nty);
    const int x0 = 1 + tx * CT, y0 = 1 + ty * CT;
    const int k0 = 1 + sg * CSEG, k1 = min(L.nz, k0 + CSEG - 1);
    const int tid = (int)threadIdx.x;
    // Every global read of step i (phi of plane i with halo 3, walls and roles of planes i - 1 and i - 2, compact
    // indices of plane i - 3) goes to registers one step ahead: the reads of step i + 1 are in flight while the
    // stages of step i run on LDS.  Element e of each table is handled by thread e (e < CBLOCK) and e + CBLOCK.
    static_assert(CPW * CPW <= 2 * CBLOCK && CNW * CNW <= CBLOCK && CXW * CXW <= CBLOCK, "tile tables per thread");
    double rphi[2];
    int8_t rbw;
    int rbr, rcr, rdm;
    auto fetch = [&](int ii) {
        for (int h = 0; h < 2; ++h) {
            const int e = tid + h * CBLOCK;
            double v = 0.0;
            if (e < CPW * CPW) {
                const int xx = x0 - 3 + e % CPW, yy = y0 - 3 + e / CPW;
                if (xx >= -3 && xx <= L.nx + 4 && yy >= -3 && yy <= L.ny + 4 && ii >= -3 && ii <= L.nz + 4)
                    v = d.phi[L.at(xx, yy, ii)];
            }
            rphi[h] = v;
        }
        rbw = 1;
        rbr = -1;
        if (tid < CNW * CNW) {
            const int xx = x0 - 2 + tid % CNW, yy = y0 - 2 + tid / CNW, jj = ii - 1;
            if (xx >= -1 && xx <= L.nx + 2 && yy >= -1 && yy <= L.ny + 2 && jj >= -1 && jj <= L.nz + 2) {
                const int o = L.at(xx, yy, jj);
                rbw = d.walls[o] == 0 ? 0 : 1;
                rbr = role[o];
            }
        }
        rcr = -1;
        if (tid < CXW * CXW) {
            const int xx = x0 - 1 + tid % CXW, yy = y0 - 1 + tid / CXW, jj = ii - 2;
            if (xx >= 0 && xx <= L.nx + 1 && yy >= 0 && yy <= L.ny + 1 && jj >= 0 && jj <= L.nz + 1)
                rcr = role[L.at(xx, yy, jj)];
        }
        rdm = -1;
        if (tid < CT * CT) {
            const int xx = x0 + tid % CT, yy = y0 + tid / CT, kk = ii - 3;
            if (xx <= L.nx && yy <= L.ny && kk >= 1 && kk <= L.nz) {
                const int o = L.at(xx, yy, kk);
                const int mm = d.map[o];
                if (d.walls[o] == 0 && mm >= 0 && mm < d.nf) rdm = mm;
            }
        }
    };
    fetch(k0 - 3);
    for (int i = k0 - 3; i <= k1 + 3; ++i) {
        {
            double* P = sphi[mod_pos(i, 3)];
            for (int h = 0; h < 2; ++h)
                if (tid + h * CBLOCK < CPW * CPW) P[tid + h * CBLOCK] = rphi[h];
            if (tid < CNW * CNW) {
                sbw[tid] = rbw;
                sbr[tid] = rbr;
            }
            if (tid < CXW * CXW) scr[tid] = rcr;
            if (tid < CT * CT) sdm[tid] = rdm;
        }
        __syncthreads();
        if (i < k1 + 3) fetch(i + 1);
        // normals of plane j = i - 1 at the non-solid nodes of [-1, n+2]^3 (k_normals), 0 elsewhere
        const int j = i - 1;
        if (j >= k0 - 2) {
            const double* P0 = sphi[mod_pos(j - 1, 3)];
            const double* P1 = sphi[mod_pos(j, 3)];
            const double* P2 = sphi[mod_pos(j + 1, 3)];
            NT(*S)[CNW * CNW] = snrm[mod_pos(j, 4)];
            NT* SC = scn[mod_pos(j, 4)];
            for (int e = tid; e < CNW * CNW; e += CBLOCK) {
                const int ex = e % CNW, ey = e / CNW;
                double cx = 0.0, cy = 0.0, cz = 0.0, cn = 0.0;
                if (sbw[e] == 0) {
                    const int c0 = (ey + 1) * CPW + (ex + 1);
                    auto A = [&](int a, int bb, int c) {
                        const double* Pz = c < 0 ? P0 : (c > 0 ? P2 : P1);
                        return Pz[c0 + a + bb * CPW];
                    };
                    cx = dx4f(A);
                    cy = dy4f(A);
                    cz = dz4f(A);
                    normalise(cx, cy, cz, cn);
                    const int r = sbr[e];
                    if (r >= 0 && cn > 1e-6)
                        wet(cx, cy, cz, fl.nw[3 * r], fl.nw[3 * r + 1], fl.nw[3 * r + 2], fl.cos_t[r], fl.sin_t[r]);
                }
                S[0][e] = (NT)cx;
                S[1][e] = (NT)cy;
                S[2][e] = (NT)cz;
                if (ex >= 2 && ex < CT + 2 && ey >= 2 && ey < CT + 2) SC[(ey - 2) * CT + (ex - 2)] = (NT)cn;
            }
        }
        __syncthreads();
        // extrapolation into the solid boundary nodes of plane j2 = i - 2 (k_normal_extrapolate)
        const int j2 = i - 2;
        if (j2 >= k0 - 1) {
            NT(*S)[CNW * CNW] = snrm[mod_pos(j2, 4)];
            for (int e = tid; e < CXW * CXW; e += CBLOCK) {
                const int r = scr[e];
                if (r > -2) continue;
                const int ex = e % CXW + 1, ey = e / CXW + 1;   // in the normal tile
                const int sidx = -r - 2;
                const unsigned mk = s1.mask[sidx];
                double ax = 0.0, ay = 0.0, az = 0.0;
#pragma unroll
                for (int q = 1; q < 19; ++q)
                    if (mk & (1u << q)) {
                        const NT(*Q)[CNW * CNW] = snrm[mod_pos(j2 + lat::EZ[q], 4)];
                        const int en = (ey + lat::EY[q]) * CNW + (ex + lat::EX[q]);
                        ax = ax + (double)Q[0][en] * lat::W[q];
                        ay = ay + (double)Q[1][en] * lat::W[q];
                        az = az + (double)Q[2][en] * lat::W[q];
                    }
                const int es = ey * CNW + ex;
                S[0][es] = (NT)(ax / s1.w[sidx]);
                S[1][es] = (NT)(ay / s1.w[sidx]);
                S[2][es] = (NT)(az / s1.w[sidx]);
            }
        }
        __syncthreads();
        // curvature force of the interior fluid nodes of plane k = i - 3 (k_curvature)
        const int k = i - 3;
        if (k >= k0 && k <= k1 && tid < CT * CT) {
            const int m = sdm[tid];
            if (m >= 0) {
                const int ex = tid % CT + 2, ey = tid / CT + 2;
                const NT(*Sm)[CNW * CNW] = snrm[mod_pos(k - 1, 4)];
                const NT(*Sc)[CNW * CNW] = snrm[mod_pos(k, 4)];
                const NT(*Sp)[CNW * CNW] = snrm[mod_pos(k + 1, 4)];
                const int c0 = ey * CNW + ex;
                auto comp = [&](int cmp) {
                    return [=](int a, int bb, int c) {
                        const NT(*Sz)[CNW * CNW] = c < 0 ? Sm : (c > 0 ? Sp : Sc);
                        return (double)Sz[cmp][c0 + a + bb * CNW];
                    };
                };
                const double curv = curvature_f(comp(0), comp(1), comp(2));
                d.force[m] = (double)Sc[0][c0];
                d.force[d.fs + m] = (double)Sc[1][c0];
                d.force[2 * d.fs + m] = (double)Sc[2][c0];
                d.force[3 * d.fs + m] = 0.5 * gamma * curv * (double)scn[mod_pos(k, 4)][tid];
            }
        }
        __syncthreads();   // the next step's first stage rewrites the per-plane tables
    }
}

// dense curvature over the whole interior, for downloads (the reference computes it everywhere)
template <class NT>
__global__ void k_curvature_dense(Dev d) {
    const Layout L = d.L;
    const long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    const long nxy = (long)L.nx * L.ny;
    if (t >= nxy * L.nz) return;
    const int k = (int)(t / nxy) + 1;
    const int r = (int)(t - (long)(k - 1) * nxy);
    const int j = r / L.nx + 1;
    const int i = r - (j - 1) * L.nx + 1;
    const int n = L.at(i, j, k);
    d.curv[n] = curvature_at<NT>(d, n);
}

// ------------------------------------------------------------------------------------------------------
// macroscopic fields and monitor reductions (compute_macro_vars Misc.F90:372-430, Monitor.F90)
// ------------------------------------------------------------------------------------------------------

// u, v, w, rho at the fluid nodes; solid nodes hold 0 as the reference's (1 - wall) factor gives
template <class T>
__global__ void k_macro(Dev d, double force_z, double* u, double* v, double* w, double* rho) {
    const int m = blockIdx.x * BLOCK + threadIdx.x;
    if (m >= d.nf) return;
    const int n = d.cpos[m];
    double ft[19];
    for (int q = 0; q < 19; ++q)
        ft[q] = sizeof(T) == 8 ? d.S.p[q][2 * m] + d.S.p[19 + q][2 * m] : (double)((const float*)d.S.p[q])[2 * m] + rest_weight(q);
    const double om = 1.0;
    rho[n] = (ft[0] + ft[1] + ft[2] + ft[3] + ft[4] + ft[5] + ft[6] + ft[7] + ft[8] + ft[9] + ft[10] + ft[11] + ft[12] +
              ft[13] + ft[14] + ft[15] + ft[16] + ft[17] + ft[18]) * om;
    const double tmp = d.force[3 * d.fs + m];
    const double fx = tmp * d.force[m];
    const double fy = tmp * d.force[d.fs + m];
    const double fz = tmp * d.force[2 * d.fs + m] + force_z;
    u[n] = (ft[1] - ft[2] + ft[7] - ft[8] + ft[9] - ft[10] + ft[11] - ft[12] + ft[13] - ft[14] - 0.5 * fx) * om;
    v[n] = (ft[3] - ft[4] + ft[7] + ft[8] - ft[9] - ft[10] + ft[15] - ft[16] + ft[17] - ft[18] - 0.5 * fy) * om;
    w[n] = (ft[5] - ft[6] + ft[11] + ft[12] - ft[13] - ft[14] + ft[15] + ft[16] - ft[17] - ft[18] - 0.5 * fz) * om;
}
// the phi = 0 inside solids side effect of compute_macro_vars (Misc.F90:423)
__global__ void k_zero_solid_phi(Dev d) {
    const Layout L = d.L;
    const long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    const long nxy = (long)L.nx * L.ny;
    if (t >= nxy * L.nz) return;
    const int k = (int)(t / nxy) + 1;
    const int r = (int)(t - (long)(k - 1) * nxy);
    const int j = r / L.nx + 1;
    const int i = r - (j - 1) * L.nx + 1;
    const int n = L.at(i, j, k);
    const int wi = d.walls[n];
    if (wi != 0) d.phi[n] = 0.0 * (double)wi + d.phi[n] * (double)(1 - wi);
}
// one block per z layer: the seven layer sums of Monitor.F90:62-84 and per-layer partials of the global
// scalars, reduced in a fixed tree (deterministic run to run)
constexpr int NPART = 12;
__global__ void __launch_bounds__(BLOCK) k_monitor_layers(Dev d, const double* u, const double* v, const double* w,
                                                           const double* rho, const double* phi_old, double* layers,
                                                           double* part) {
    const Layout L = d.L;
    const int k = blockIdx.x + 1;
    double s[NPART];
    for (int m = 0; m < NPART; ++m) s[m] = 0.0;
    double umax = 0.0, dphi = 0.0;
    const int nxy = L.nx * L.ny;
    for (int r = threadIdx.x; r < nxy; r += BLOCK) {
        const int j = r / L.nx + 1, i = r % L.nx + 1;
        const int n = L.at(i, j, k);
        const int wi = d.walls[n];
        const double om = (double)(1 - wi);
        const double ph = d.phi[n];
        const double temp = (u[n] * u[n] + v[n] * v[n] + w[n] * w[n]) * om;
        umax = fmax(umax, temp);
        if (ph > 0.999) s[7] += temp;
        else if (ph < -0.999) s[8] += temp;
        s[2] += 0.5 * (1.0 + ph) * om;
        s[3] += 0.5 * (1.0 - ph) * om;
        s[4] += rho[n] * 0.5 * (1.0 + ph) * om;
        s[5] += rho[n] * 0.5 * (1.0 - ph) * om;
        s[0] += w[n] * 0.5 * (1.0 + ph) * om;
        s[1] += w[n] * 0.5 * (1.0 - ph) * om;
        s[6] += rho[n] * om;
        if (wi == 0) {
            if (ph < -0.99) { s[9] += rho[n]; s[10] += 1.0; }
            if (ph > 0.99) { s[11] += rho[n]; }
            dphi = fmax(dphi, fabs(ph - phi_old[n]) * om);
        }
    }
    // count of fluid-1 nodes (phi > 0.99) for the capillary-pressure monitor
    double c1 = 0.0;
    for (int r = threadIdx.x; r < nxy; r += BLOCK) {
        const int n = L.at(r % L.nx + 1, r / L.nx + 1, k);
        if (d.walls[n] == 0 && d.phi[n] > 0.99) c1 += 1.0;
    }
    // wavefront reductions (a fixed butterfly over the 64 lanes), then the four waves' results in wave order
    constexpr int NQ = NPART + 3, NW = BLOCK / 64;
    double q15[NQ];
    for (int m = 0; m < NPART; ++m) q15[m] = s[m];
    q15[NPART] = umax;
    q15[NPART + 1] = dphi;
    q15[NPART + 2] = c1;
#pragma unroll
    for (int m = 0; m < NQ; ++m)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double o = __shfl_xor(q15[m], off, 64);
            q15[m] = (m == NPART || m == NPART + 1) ? fmax(q15[m], o) : q15[m] + o;
        }
    __shared__ double sh[NQ][NW];
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int m = 0; m < NQ; ++m) sh[m][wave] = q15[m];
    __syncthreads();
    double out[NQ];
    if (threadIdx.x == 0)
        for (int m = 0; m < NQ; ++m) {
            double a = sh[m][0];
            for (int w = 1; w < NW; ++w) a = (m == NPART || m == NPART + 1) ? fmax(a, sh[m][w]) : a + sh[m][w];
            out[m] = a;
        }
    if (threadIdx.x == 0) {
        for (int m = 0; m < 7; ++m) layers[m * L.nz + (k - 1)] = out[m];
        for (int m = 7; m < NPART + 3; ++m) part[(m - 7) * L.nz + (k - 1)] = out[m];
    }
}
// Initial state of a new run at rest, on the device (initialization_new_multi / _pdf, Init_multiphase.F90:
// 243-470, the same expressions as mflbm_init_fields with u = 0): phi by initial_fluid_distribution_option
// over the halo-4 box, rho1 = (1 + phi)/2 and rho2 = (1 - phi)/2, and the equilibrium populations of both
// colours over the halo-1 box into the dense staging arrays (then gathered like an upload).
__global__ void k_init_dense(Dev d, double* __restrict__ dense, long ps, int option, double z0, double phi_inlet,
                             int open_z) {
    const Layout L = d.L;
    const long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    const long nb = (long)(L.nx + 8) * (L.ny + 8) * (L.nz + 8);
    if (t >= nb) return;
    const int i = (int)(t % (L.nx + 8)) - 3, j = (int)((t / (L.nx + 8)) % (L.ny + 8)) - 3,
              k = (int)(t / ((long)(L.nx + 8) * (L.ny + 8))) - 3;
    const double x = i, y = j, z = k;
    double v;
    if (option == 1) {
        v = -1.0;
        if (z <= z0) v = 1.0;
    } else if (option == 2) {
        v = 1.0;
        if (z <= z0) v = -1.0;
    } else if (option == 3 || option == 4) {
        const double dx = x - (L.nx + 1) * 0.0, dz = z - (L.nz + 1) * 0.5, dy = y - (L.ny + 1) * 0.5;
        const bool in = dx * dx + dz * dz + dy * dy <= z0 * z0;
        v = option == 3 ? (in ? 1.0 : -1.0) : (in ? -1.0 : 1.0);
    } else {
        const double dx = x - (L.nx + 1) * 0.5, dz = z - (L.nz + 1) * 0.5, dy = y - (L.ny + 1) * 0.5;
        v = dx * dx + dz * dz + dy * dy <= z0 * z0 ? 1.0 : -1.0;
    }
    if (open_z && z <= 0) v = phi_inlet;
    const int o = L.at(i, j, k);
    d.phi[o] = v;
    if (i < 0 || i > L.nx + 1 || j < 0 || j > L.ny + 1 || k < 0 || k > L.nz + 1) return;
    const double rho = 1.0, usq = 0.0, eu = 0.0, uu = 0.0;
    const double r1 = rho * (1.0 + v) * 0.5, r2 = rho * (1.0 - v) * 0.5;
    for (int q = 0; q < 19; ++q) {
        const double wq = lat::W[q];
        double a, b;
        if (q == 0) {
            a = r1 * wq + r1 * wq * (-1.5 * usq);
            b = r2 * wq + r2 * wq * (-1.5 * usq);
        } else {
            a = r1 * wq + r1 * wq * (3.0 * eu + 4.5 * uu * uu - 1.5 * usq);
            b = r2 * wq + r2 * wq * (3.0 * eu + 4.5 * uu * uu - 1.5 * usq);
        }
        dense[q * ps + o] = a;
        dense[(19 + q) * ps + o] = b;
    }
}
// convective-outlet history of a new run: the populations and phi of the outlet layer z = nz
__global__ void k_init_convec(Dev d, const double* __restrict__ dense, long ps) {
    const Layout L = d.L;
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= (L.nx + 2) * (L.ny + 2)) return;
    const int i = t % (L.nx + 2), j = t / (L.nx + 2);
    const long pl = (long)(L.nx + 2) * (L.ny + 2);
    const int o = L.at(i, j, L.nz);
    for (int q = 0; q < 19; ++q) {
        d.cf[t + q * pl] = dense[q * ps + o];
        d.cg[t + q * pl] = dense[(19 + q) * ps + o];
    }
    d.cphi[t] = d.phi[o];
}

__global__ void k_copy_phi_old(Dev d, double* phi_old) {
    const long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (t >= d.L.total) return;
    phi_old[t] = d.phi[t];
}

// ------------------------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------------------------



}  // namespace

struct mflbm_ctx {
    mflbm_params p{};
    Layout L{};
    Dev d{};
    hipStream_t stream = nullptr;
    int device = 0;
    long ps = 0;                // stride of the dense arrays
    double* dense_pdf = nullptr;   // 38 dense arrays: upload/download staging, keeps never-touched slots
    double* fields = nullptr;   // phi, the normal records (4 arrays' worth), curv, u, v, w, rho, phi_old
    float* fields32 = nullptr;  // FP32 storage modes: the normal records in FP32
    int* role = nullptr;        // dense: fluid boundary index (>= 0) or -(extrapolation list index) - 2, else -1
    unsigned char* qbuf = nullptr;   // brick flags: notm, notp (two buffers each), bstatic, nzb, nzd, hasb
    int* cbuf = nullptr;             // chain lists: NWORK work lists (NXCD sub-lists of lcap int4 items each), the
                                     // brick offsets of the 3 sorted lists (nbricks + 1 each), 2 sets of NWORK x NXCD
                                     // counters
    int lcap = 0;
    double* phic = nullptr;          // compact phi (Dev::phic)
    bool phi_stale = false;          // the dense phi at fluid nodes may lag the compact phi (outside the chain's bricks)
    long nbricks = 0;
    int band_cur = 0;                // band buffers the lattice update fills until the next chain
    bool full_chain = true;          // the next chain evaluates every brick (new geometry or state)
    bool solid_zeroed = false;       // the macroscopic pass zeroed phi in solids since the last chain
    Fold fold;                       // periodic axes folded into the neighbour map
    bool stepped = false;            // a step ran since the state was uploaded or built
    bool chain_fused = false;   // one fused pass for normals, extrapolation and curvature (MFLBM_CHAIN=fused)
    bool chain_all = false;     // the chain over the whole lists, without the brick plan (MFLBM_CHAIN=all)
    int8_t* walls = nullptr;
    double* planes = nullptr;   // w_in, cphi, cf(19), cg(19)
    double* u = nullptr;
    double* v = nullptr;
    double* w = nullptr;
    double* rho = nullptr;
    double* phi_old = nullptr;
    double* red = nullptr;      // monitor layer sums and partials
    // geometry-dependent device buffers
    void* compact = nullptr;    // 38 population slices (compact slots + link slots), FP64 or FP32
    bool fp32 = false;          // precision 1 or 2: FP32 storage (s, a) of collide.h
    double* force = nullptr;    // 4 compact force arrays
    int* map = nullptr;         // Dev::map, then Dev::pmap (2 x L.total)
    int* ibuf = nullptr;        // cpos, normal list, solid lists, fluid list positions
    int* wbuf = nullptr;        // per normal-list node: index into the fluid boundary list or -1
    unsigned* mbuf = nullptr;   // solid list masks
    double* dbuf = nullptr;     // solid weights, fluid list normals / cos / sin
    unsigned long long* wmask = nullptr;  // per wave and direction: fluid nodes with a link
    int* wbase = nullptr;       // per wave of fluid nodes: link counts before it
    SolidList sl{}, sl1{};
    FluidList fl{};
    NodeList nrm{};
    bool have_geom = false, have_state = false;
    bool timing = false;
    std::vector<hipEvent_t> ev;   // pool of 4 events per timed step: start, after collide, before chain, end
    size_t ev_used = 0;
    double t_sum[3] = {0, 0, 0};
    long t_cnt[3] = {0, 0, 0};
    std::string err;
};

namespace {

CollideParams collide_params(const mflbm_params& p) {
    return CollideParams{p.gamma, p.beta, p.force_z, p.la_nui1, p.la_nui2};
}

bool open_z(const mflbm_params& p) { return p.kper == 0 && p.wall_z_min == 0 && p.wall_z_max == 0; }
// The periodic axis folded into the neighbour map (layout.h): the one periodic axis of a domain periodic along z
// only, or along y only without an inlet/outlet face (which addresses the y halo rows); never with a porous
// plate, whose planes address the halo rows
Fold fold_of(const mflbm_params& p) {
    const bool plate = p.porous_plate_cmd == 1 || p.porous_plate_cmd == 2;
    return Fold{p.jper == 1 && p.kper == 0 && !open_z(p) && !plate, p.kper == 1 && p.jper == 0 && !plate};
}


// work list of one chain stage (0 normals, 1 normal extrapolation, 2 phi extrapolation) and its counter in set `set`
BrickWork brick_work(const mflbm_ctx* c, int work, int set) {
    const long nb = c->nbricks;
    const size_t lw = (size_t)NXCD * c->lcap;   // items per work list
    const int4* lists = (const int4*)c->cbuf;
    return BrickWork{lists + (size_t)work * lw,
                     c->cbuf + 4 * NWORK * lw + 3 * (size_t)(nb + 1) + NXCD * (NWORK * set + work), c->lcap};
}
#define CHAIN_LAUNCH(c, kern, ...)                                                                           \
    do {                                                                                                     \
        hipLaunchKernelGGL(kern, dim3(CHAIN_BLOCKS), dim3(BLOCK), 0, (c)->stream, __VA_ARGS__);              \
        const hipError_t e_ = hipGetLastError();                                                             \
        if (e_ != hipSuccess) return fail((c), MFLBM_ERR_HIP, "launch %s: %s", #kern, hipGetErrorString(e_));  \
    } while (0)

// the normal field over the bricks of the work lists of set `set`
template <class NT>
int launch_normals_listed(mflbm_ctx* c, int set) {
    CHAIN_LAUNCH(c, k_normals<NT>, c->d, c->nrm, (const int*)c->wbuf, c->fl, brick_work(c, 0, set));
    CHAIN_LAUNCH(c, k_normal_extrapolate<NT>, c->d, c->sl1, brick_work(c, 1, set));
    return MFLBM_OK;
}

// the dense phi at every fluid node (the lattice update writes the compact phi; the chain keeps no dense copy)
int ensure_dense_phi(mflbm_ctx* c) {
    if (c->phi_stale) LAUNCH(c, k_phi_dense, nblocks(c->d.nf), c->d);
    c->phi_stale = false;
    return MFLBM_OK;
}

// Open inlet and outlet conditions a step hands to the chain's plan launch (kind 0: none)
struct FaceWork {
    int in = 0, out = 0;
    bool after = false;
    FaceArgs a{};
    int gface = 0;
};
template <class T, int IN, int OUT, bool AFTER>
int launch_plan_k(mflbm_ctx* c, const ChainPlan& p, int full, const FaceWork& f) {
    const int ntx = (c->d.nbx + TBX - 1) / TBX, nty = (c->d.nby + TBY - 1) / TBY, ntz = (c->d.nbz + TBZ - 1) / TBZ;
    const int nface = IN != 0 ? 2 * f.gface : 0;
    hipLaunchKernelGGL((k_brick_plan<T, IN, OUT, AFTER>), dim3(nface + ntx * nty * ntz), dim3(BLOCK), 0, c->stream,
                       c->d, p, full, f.a, f.gface, ntx, nty);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(c, MFLBM_ERR_HIP, "launch k_brick_plan: %s", hipGetErrorString(e));
    return MFLBM_OK;
}
template <class T>
int launch_plan(mflbm_ctx* c, const ChainPlan& p, int full, const FaceWork& f) {
#define MFLBM_PLAN(IN, OUT)                                                                                     \
    return f.after ? launch_plan_k<T, IN, OUT, true>(c, p, full, f) : launch_plan_k<T, IN, OUT, false>(c, p, full, f);
    if (f.in == 1 && f.out == 1) { MFLBM_PLAN(1, 1) }
    if (f.in == 1 && f.out == 2) { MFLBM_PLAN(1, 2) }
    if (f.in == 2 && f.out == 1) { MFLBM_PLAN(2, 1) }
    if (f.in == 2 && f.out == 2) { MFLBM_PLAN(2, 2) }
#undef MFLBM_PLAN
    return launch_plan_k<double, 0, 0, false>(c, p, full, f);
}

// The colour-gradient chain of one step (Phase_gradient.F90:5-204): the brick plan, then phi extrapolation,
// normals and normal extrapolation over the bricks it listed (or the fused chain).  The plan reads the band
// flags the last lattice update filled and clears the other buffers, which the next lattice update fills.
template <class NT>
int launch_gradient_t(mflbm_ctx* c, const FaceWork& faces) {
    if (c->chain_all) {   // every entry of every list, from the dense phi; nzd cleared and set again by the normals
        const int nb = (int)c->nbricks;
        if (c->phi_stale) LAUNCH(c, k_phi_dense_clear, nblocks(std::max(c->d.nf, nb)), c->d, c->d.nzd, nb);
        else LAUNCH(c, k_clear_bytes, nblocks(nb), c->d.nzd, nb);
        c->phi_stale = false;
        LAUNCH(c, k_phi_extrapolate, NXCD * xcd_per(c->sl.n), c->d, c->sl, xcd_per(c->sl.n));
        LAUNCH(c, k_normals_all<NT>, nblocks(c->nrm.n), c->d, c->nrm, (const int*)c->wbuf, c->fl);
        LAUNCH(c, k_normal_extrapolate_all<NT>, nblocks(c->sl1.n), c->d, c->sl1);
        c->full_chain = false;
        c->solid_zeroed = false;
        return MFLBM_OK;
    }
    const long nb = c->nbricks;
    const int set = c->band_cur;
    {
        ChainPlan p;
        p.notm = c->qbuf + (size_t)set * nb;
        p.notp = c->qbuf + (size_t)(2 + set) * nb;
        p.next_m = c->qbuf + (size_t)(set ^ 1) * nb;
        p.next_p = c->qbuf + (size_t)(2 + (set ^ 1)) * nb;
        p.hasb = c->qbuf + 7 * (size_t)nb;
        const size_t lw = (size_t)NXCD * c->lcap;
        const int* offs = c->cbuf + 4 * NWORK * lw;
        p.off_n = offs;
        p.off_s = offs + (nb + 1);
        p.off_1 = offs + 2 * (nb + 1);
        int4* lists = (int4*)c->cbuf;
        for (int l = 0; l < NWORK; ++l) p.list[l] = lists + (size_t)l * lw;
        p.cap = c->lcap;
        int* cnts = c->cbuf + 4 * NWORK * lw + 3 * (size_t)(nb + 1);
        p.cnt = cnts + NWORK * NXCD * set;
        p.cnt_next = cnts + NWORK * NXCD * (set ^ 1);
        // no brick quiet after new geometry or state, and none with the fused chain, which does not use the flags
        const int full = (c->full_chain || c->chain_fused) ? 1 : 0;
        const int rp = c->fp32 ? launch_plan<float>(c, p, full, faces) : launch_plan<double>(c, p, full, faces);
        if (rp) return rp;
        c->band_cur ^= 1;
        c->d.notm = p.next_m;
        c->d.notp = p.next_p;
        c->full_chain = false;
        c->solid_zeroed = false;
    }
    CHAIN_LAUNCH(c, k_phi_extrapolate_b, c->d, c->sl, brick_work(c, 2, set));
    if (c->chain_fused) {   // (its LDS planes are loaded from the dense phi)
        const int re = ensure_dense_phi(c);
        if (re) return re;
        const Layout& L = c->L;
        const int ntx = (L.nx + CT - 1) / CT, nty = (L.ny + CT - 1) / CT, nseg = (L.nz + CSEG - 1) / CSEG;
        const long nblk = (long)ntx * nty * nseg;
        const int per = (int)((nblk + NXCD - 1) / NXCD);
        hipLaunchKernelGGL(k_chain_force<NT>, dim3(NXCD * per), dim3(CBLOCK), 0, c->stream, c->d, c->fl, c->sl1,
                           (const int*)c->role, c->p.gamma, ntx, nty, per);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return fail(c, MFLBM_ERR_HIP, "launch k_chain_force: %s", hipGetErrorString(e));
        return MFLBM_OK;
    }
    return launch_normals_listed<NT>(c, set);
}
int launch_gradient(mflbm_ctx* c, const FaceWork& faces = FaceWork{}) {
    return c->fp32 ? launch_gradient_t<float>(c, faces) : launch_gradient_t<double>(c, faces);
}

template <class T>
int launch_faces(mflbm_ctx* c, bool odd, FaceWork* defer);

int launch_step(mflbm_ctx* c, int ntime) {
    c->stepped = true;
    c->phi_stale = true;
    const mflbm_params& p = c->p;
    const CollideParams P = collide_params(p);
    const bool odd = (ntime % 2) != 0;
    const bool t = c->timing;
    hipEvent_t* ev = nullptr;
    if (t) {
        if (c->ev_used + 4 > c->ev.size()) {
            for (int m = 0; m < 4; ++m) {
                hipEvent_t e;
                HIPCHK(c, hipEventCreate(&e));
                c->ev.push_back(e);
            }
        }
        ev = &c->ev[c->ev_used];
        c->ev_used += 4;
        HIPCHK(c, hipEventRecord(ev[0], c->stream));
    }
    const int per = xcd_per(c->d.nf);
#define MFLBM_COLLIDE32(C_, PP)                                                                                  \
    if (c->chain_fused) {                                                                                        \
        if (odd) LAUNCH(c, (k_collide32<true, C_, false>), NXCD * per, c->d, PP, p.gamma, per);                   \
        else LAUNCH(c, (k_collide32<false, C_, false>), NXCD * per, c->d, PP, p.gamma, per);                      \
    } else {                                                                                                     \
        if (odd) LAUNCH(c, (k_collide32<true, C_, true>), NXCD * per, c->d, PP, p.gamma, per);                    \
        else LAUNCH(c, (k_collide32<false, C_, true>), NXCD * per, c->d, PP, p.gamma, per);                       \
    }
    if (c->fp32 && c->p.precision == 2) {
        const CollideParamsDev<double> PD{p.gamma, p.beta, p.force_z, p.la_nui1, p.la_nui2};
        MFLBM_COLLIDE32(double, PD)
    } else if (c->fp32) {
        const CollideParamsDev<float> P32{(float)p.gamma, (float)p.beta, (float)p.force_z, (float)p.la_nui1, (float)p.la_nui2};
        MFLBM_COLLIDE32(float, P32)
    } else if (c->chain_fused) {
        if (odd) LAUNCH(c, (k_collide<true, false>), NXCD * per, c->d, P, per);
        else LAUNCH(c, (k_collide<false, false>), NXCD * per, c->d, P, per);
    } else {
        if (odd) LAUNCH(c, (k_collide<true, true>), NXCD * per, c->d, P, per);
        else LAUNCH(c, (k_collide<false, true>), NXCD * per, c->d, P, per);
    }
#undef MFLBM_COLLIDE32
    if (t) HIPCHK(c, hipEventRecord(ev[1], c->stream));
    FaceWork faces;
    FaceWork* defer = c->chain_all ? nullptr : &faces;   // (no plan launch to take the faces)
    const int rc = c->fp32 ? launch_faces<float>(c, odd, defer) : launch_faces<double>(c, odd, defer);
    if (rc) return rc;
    if (t) HIPCHK(c, hipEventRecord(ev[2], c->stream));
    const int rg = launch_gradient(c, faces);
    if (rg) return rg;
    if (t) HIPCHK(c, hipEventRecord(ev[3], c->stream));
    return MFLBM_OK;
}

// periodic wrap and boundary conditions of one step, after the lattice update.  The combined inlet + outlet
// condition of a domain without a porous plate is handed to the chain's plan launch through `defer` (the chain
// follows at once), where it overlaps the plan; otherwise it is launched here.
template <class T>
int launch_faces(mflbm_ctx* c, bool odd, FaceWork* defer) {
    const mflbm_params& p = c->p;
    const Layout& L = c->L;
    // periodic wrap (the reference's MPI exchange of one rank with itself)
    if (p.kper == 1 || p.jper == 1) {
        // population exchange of the axes not folded into the neighbour map
        const Fold F = fold_of(p);
        const bool wz = p.kper == 1 && !F.z, wy = p.jper == 1 && !F.y, we = p.kper == 1 && p.jper == 1 && !(F.y && F.z);
        if (!odd) {
            if (wz) LAUNCH(c, k_wrap_pull_z<T>, nblocks((long)L.nx * L.ny), c->d);
            if (wy) LAUNCH(c, k_wrap_pull_y<T>, nblocks((long)L.nx * L.nz), c->d);
            if (we) LAUNCH(c, k_wrap_pull_edge<T>, nblocks(L.nx), c->d);
        } else {
            if (wz) LAUNCH(c, k_wrap_push_z<T>, nblocks((long)L.nx * L.ny), c->d);
            if (wy) LAUNCH(c, k_wrap_push_y<T>, nblocks((long)L.nx * L.nz), c->d);
            if (we) LAUNCH(c, k_wrap_push_edge<T>, nblocks(L.nx), c->d);
        }
        LAUNCH(c, k_wrap_phi, nblocks((long)L.nx * (L.ny + 8) * (L.nz + 8)), c->d, p.jper, p.kper);
    }
    const unsigned gface = nblocks((long)L.nx * L.ny);
    const bool both = (p.inlet_bc == 1 || p.inlet_bc == 2) && (p.outlet_bc == 1 || p.outlet_bc == 2);
    if (open_z(p) && both && L.nz >= 8) {
        const FaceArgs a{p.phi_inlet, p.sa_inject, p.relaxation, p.rho_in, p.uin_avg, p.rho_out};
        const int g = (int)gface;
        const bool plate = p.porous_plate_cmd == 1 || p.porous_plate_cmd == 2;
        if (defer && !plate) {
            defer->in = p.inlet_bc;
            defer->out = p.outlet_bc;
            defer->after = odd;
            defer->a = a;
            defer->gface = g;
            return MFLBM_OK;
        }
#define MFLBM_FACES(IN, OUT)                                                                                    \
    if (odd) LAUNCH(c, (k_open_faces<T, IN, OUT, true>), 2 * gface, c->d, a, g);                                \
    else LAUNCH(c, (k_open_faces<T, IN, OUT, false>), 2 * gface, c->d, a, g);
        if (p.inlet_bc == 1 && p.outlet_bc == 1) { MFLBM_FACES(1, 1) }
        else if (p.inlet_bc == 1) { MFLBM_FACES(1, 2) }
        else if (p.outlet_bc == 1) { MFLBM_FACES(2, 1) }
        else { MFLBM_FACES(2, 2) }
#undef MFLBM_FACES
    } else if (open_z(p)) {
        if (!odd) {
            if (p.inlet_bc == 1) LAUNCH(c, k_inlet_velocity_before<T>, gface, c->d, p.phi_inlet, p.sa_inject, p.relaxation);
            else if (p.inlet_bc == 2) LAUNCH(c, k_inlet_pressure_before<T>, gface, c->d, p.phi_inlet, p.sa_inject, p.rho_in, p.relaxation);
            if (p.outlet_bc == 1) LAUNCH(c, k_outlet_convective_before<T>, gface, c->d, p.uin_avg);
            else if (p.outlet_bc == 2) LAUNCH(c, k_outlet_pressure_before<T>, gface, c->d, p.rho_out);
        } else {
            if (p.inlet_bc == 1) LAUNCH(c, k_inlet_velocity_after<T>, gface, c->d, p.phi_inlet, p.sa_inject, p.relaxation);
            else if (p.inlet_bc == 2) LAUNCH(c, k_inlet_pressure_after<T>, gface, c->d, p.phi_inlet, p.sa_inject, p.rho_in, p.relaxation);
            if (p.outlet_bc == 1) LAUNCH(c, k_outlet_convective_after<T>, gface, c->d, p.uin_avg);
            else if (p.outlet_bc == 2) LAUNCH(c, k_outlet_pressure_after<T>, gface, c->d, p.rho_out);
        }
    }
    if (p.porous_plate_cmd == 1 || p.porous_plate_cmd == 2) {
        const int zp = p.z_porous_plate;
        if (zp >= 1 && zp <= L.nz) LAUNCH(c, k_porous_plate<T>, gface, c->d, zp, p.porous_plate_cmd - 1, odd ? 1 : 0);
    }
    return MFLBM_OK;
}

// dense staging -> compact storage (upload, device initialisation) and back (download)
template <class T>
int gather_state(mflbm_ctx* c) {
    LAUNCH(c, k_gather<T>, nblocks(c->d.nc), c->d, (const double*)c->dense_pdf, c->ps);
    LAUNCH(c, (k_links<false, T>), nblocks(c->d.nf), c->d, c->dense_pdf, c->ps);
    if (c->fold.y || c->fold.z)
        LAUNCH(c, k_fold_upload<T>, fold_blocks(c->L), c->d, c->fold, (const double*)c->dense_pdf, c->ps);
    LAUNCH(c, k_phic_gather, nblocks(c->d.nf), c->d);
    c->phi_stale = false;
    c->stepped = false;
    return MFLBM_OK;
}
template <class T>
int scatter_state(mflbm_ctx* c) {
    LAUNCH(c, k_scatter<T>, nblocks(c->d.nc), c->d, c->dense_pdf, c->ps);
    LAUNCH(c, (k_links<true, T>), nblocks(c->d.nf), c->d, c->dense_pdf, c->ps);
    return MFLBM_OK;
}

void free_geometry(mflbm_ctx* c) {
    void* bufs[] = {c->compact, c->force, c->map, c->ibuf, c->wbuf, c->mbuf, c->dbuf, c->wmask, c->wbase, c->role, c->qbuf, c->cbuf, c->phic};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    c->compact = nullptr;
    c->force = nullptr;
    c->map = nullptr;
    c->ibuf = nullptr;
    c->wbuf = nullptr;
    c->mbuf = nullptr;
    c->dbuf = nullptr;
    c->wmask = nullptr;
    c->wbase = nullptr;
    c->role = nullptr;
    c->qbuf = nullptr;
    c->cbuf = nullptr;
    c->phic = nullptr;
    c->d.phic = nullptr;
    c->have_geom = false;
    c->have_state = false;
}

}  // namespace

extern "C" const char* mflbm_create_error(void) { return g_create_error.c_str(); }
extern "C" const char* mflbm_last_error(const mflbm_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

extern "C" int mflbm_create(const mflbm_params* prm, mflbm_ctx** out) {
    if (!out) return fail(nullptr, MFLBM_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!prm) return fail(nullptr, MFLBM_ERR_ARG, "params is NULL");
    if (prm->nx < 3 || prm->ny < 3 || prm->nz < 3) return fail(nullptr, MFLBM_ERR_ARG, "lattice must be at least 3^3");
    if (prm->wall_x_min != 1 || prm->wall_x_max != 1)
        return fail(nullptr, MFLBM_ERR_ARG, "x needs solid walls at both ends (the reference rejects anything else)");
    if (prm->precision < 0 || prm->precision > 2)
        return fail(nullptr, MFLBM_ERR_ARG, "precision must be 0 (FP64), 1 (FP32) or 2 (FP32 storage, FP64 collision), got %d",
                    prm->precision);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(nullptr, MFLBM_ERR_NODEV, "no HIP device visible");
    mflbm_ctx* c = new (std::nothrow) mflbm_ctx();
    if (!c) return fail(nullptr, MFLBM_ERR_NOMEM, "context allocation failed");
    c->p = *prm;
    c->fp32 = prm->precision != 0;
    {
        // the fused chain is bit-identical but evaluates every node, so it is slower than the brick lists on the
        // 256^3 benchmark although the lattice update it feeds is 8 % faster: the list kernels stay the default.
        // FP32 collision: its rounding leaves phi outside the band at nearly every node, so the plan would list
        // nearly every brick and the whole-list chain is the faster (MFLBM_CHAIN=bricks keeps the plan)
        const char* e = std::getenv("MFLBM_CHAIN");
        const std::string m = e ? e : "";
        c->chain_fused = m == "fused";
        c->chain_all = m == "all" || (m.empty() && prm->precision == 1);
    }
    Layout& L = c->L;
    L = make_layout(prm->nx, prm->ny, prm->nz);
    if (L.total > (1L << 31) - 1) {
        delete c;
        return fail(nullptr, MFLBM_ERR_ARG, "lattice too large for 32-bit node offsets");
    }
    const long ps = ((L.total + 31) / 32) * 32;   // 256-byte aligned array stride
    c->ps = ps;
    auto bail = [&](int code) {
        std::string m = c->err;
        mflbm_destroy(c);
        return fail(nullptr, code, "%s", m.c_str());
    };
    if (hipGetDevice(&c->device) != hipSuccess) return bail(fail(c, MFLBM_ERR_HIP, "hipGetDevice failed"));
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(c, MFLBM_ERR_HIP, "hipStreamCreate failed"));
    const long plane = (long)(L.nx + 2) * (L.ny + 2);
    if (hipMalloc(&c->dense_pdf, sizeof(double) * ps * 38) != hipSuccess ||
        hipMalloc(&c->fields, sizeof(double) * ps * 11) != hipSuccess ||
        hipMalloc(&c->walls, (size_t)ps) != hipSuccess ||
        hipMalloc(&c->planes, sizeof(double) * plane * 40) != hipSuccess ||
        hipMalloc(&c->red, sizeof(double) * (size_t)(L.nz + 1) * 32) != hipSuccess)
        return bail(fail(c, MFLBM_ERR_NOMEM, "device allocation of %.2f GB failed",
                         (38.0 + 11.0) * 8.0 * ps / 1e9));
    if (hipMemset(c->dense_pdf, 0, sizeof(double) * ps * 38) != hipSuccess ||
        hipMemset(c->fields, 0, sizeof(double) * ps * 11) != hipSuccess ||
        hipMemset(c->walls, 1, (size_t)ps) != hipSuccess || hipMemset(c->planes, 0, sizeof(double) * plane * 40) != hipSuccess)
        return bail(fail(c, MFLBM_ERR_HIP, "hipMemset failed"));
    if (c->fp32 && (hipMalloc(&c->fields32, sizeof(float) * ps * 4) != hipSuccess ||
                    hipMemset(c->fields32, 0, sizeof(float) * ps * 4) != hipSuccess))
        return bail(fail(c, MFLBM_ERR_NOMEM, "device allocation of the FP32 normal field failed"));
    Dev& d = c->d;
    if (c->fp32) {
        d.cn32 = (float4*)c->fields32;
    }
    d.L = L;
    d.phi = c->fields;
    d.cn = (double4*)(c->fields + ps);
    d.curv = c->fields + 5 * ps;
    c->u = c->fields + 6 * ps;
    c->v = c->fields + 7 * ps;
    c->w = c->fields + 8 * ps;
    c->rho = c->fields + 9 * ps;
    c->phi_old = c->fields + 10 * ps;
    d.walls = c->walls;
    d.w_in = c->planes;
    d.cphi = c->planes + plane;
    d.cf = c->planes + 2 * plane;
    d.cg = c->planes + 21 * plane;
    if (hipDeviceSynchronize() != hipSuccess) return bail(fail(c, MFLBM_ERR_HIP, "hipDeviceSynchronize failed"));
    *out = c;
    return MFLBM_OK;
}

extern "C" void mflbm_destroy(mflbm_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto e : c->ev)
        if (e) (void)hipEventDestroy(e);
    free_geometry(c);
    void* bufs[] = {c->dense_pdf, c->fields, c->fields32, c->walls, c->planes, c->red};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" int mflbm_set_scalars(mflbm_ctx* c, double force_z, double rho_in, double uin_avg, double relaxation) {
    if (!c) return MFLBM_ERR_ARG;
    c->p.force_z = force_z;
    c->p.rho_in = rho_in;
    c->p.uin_avg = uin_avg;
    c->p.relaxation = relaxation;
    return MFLBM_OK;
}

extern "C" int mflbm_upload_geometry(mflbm_ctx* c, const int8_t* walls_h2, int32_t ns, const int32_t* s_ijk,
                                     const int32_t* s_nfl, const int32_t* s_nbr, const double* s_w, int32_t nf,
                                     const int32_t* f_ijk, const double* f_nw, const double* f_theta, const double* w_in) {
    if (!c) return MFLBM_ERR_ARG;
    if (!walls_h2 || ns < 0 || nf < 0 || (ns > 0 && (!s_ijk || !s_nfl || !s_nbr || !s_w)) ||
        (nf > 0 && (!f_ijk || !f_nw || !f_theta)))
        return fail(c, MFLBM_ERR_ARG, "missing geometry array");
    const Layout& L = c->L;
    const mflbm_params& p = c->p;
    // bounds of the list coordinates against the ranges the kernels index (phi halo 4, normal halo 2)
    std::vector<unsigned> smask((size_t)ns, 0u);
    for (int t = 0; t < ns; ++t) {
        const int i = s_ijk[3 * t], j = s_ijk[3 * t + 1], k = s_ijk[3 * t + 2];
        if (i < -2 || i > L.nx + 3 || j < -2 || j > L.ny + 3 || k < -2 || k > L.nz + 3 || s_nfl[t] < 0 || s_nfl[t] > 18)
            return fail(c, MFLBM_ERR_ARG, "solid boundary node %d out of range", t);
        int prev = 0;
        for (int m = 0; m < s_nfl[t]; ++m) {
            const int q = s_nbr[18 * t + m];
            // the reference builds neighbour_list in increasing direction order; the bit mask keeps that order
            if (q <= prev || q > 18) return fail(c, MFLBM_ERR_ARG, "solid boundary node %d: neighbour list not increasing in 1..18", t);
            smask[t] |= 1u << q;
            prev = q;
        }
    }
    for (int t = 0; t < nf; ++t) {
        const int i = f_ijk[3 * t], j = f_ijk[3 * t + 1], k = f_ijk[3 * t + 2];
        if (i < -1 || i > L.nx + 2 || j < -1 || j > L.ny + 2 || k < -1 || k > L.nz + 2)
            return fail(c, MFLBM_ERR_ARG, "fluid boundary node %d out of range", t);
    }
    std::vector<int8_t> wp(L.total, 1);
    pack_host(L, 2, walls_h2, wp.data());
    // compact node set: interior fluid nodes, then the support nodes (see device_common.h)
    std::vector<int> extra_z;
    if ((p.porous_plate_cmd == 1 || p.porous_plate_cmd == 2) && p.z_porous_plate >= 1 && p.z_porous_plate <= L.nz)
        extra_z = {p.z_porous_plate - 1, p.z_porous_plate, p.z_porous_plate + 1};
    const TileOrder T{tile_rows(), L};
    CompactSet cset;
    build_compact(L, T, wp, open_z(p), p.kper == 1, p.jper == 1, extra_z, cset, fold_of(p));
    c->fold = cset.fold;
    const std::vector<int>& cidx = cset.cidx;
    const std::vector<int>& cpos = cset.cpos;
    const int nfl = cset.nfl;
    LinkSet links;
    build_links(L, cset, links);
    const std::vector<int>& wbase = links.wbase;
    const long* nlink = links.nlink;
    const int nc = (int)cpos.size();
    // nodes whose normal is computed: walls == 0 in [-1, n+2]^3
    std::vector<int> nrm;
    T.visit(-1, L.nx + 2, -1, L.ny + 2, -1, L.nz + 2, [&](int i, int j, int k) {
        if (wp[L.at(i, j, k)] == 0) nrm.push_back(L.at(i, j, k));
    });
    // The chain lists are sorted by brick (brick_of), in the tile order within a brick;
    // every entry is independent, so the order changes no result
    const int nbx = (L.px + (1 << BRICK_SX) - 1) >> BRICK_SX, nby = (L.ny + 8 + (1 << BRICK_SY) - 1) >> BRICK_SY,
              nbz = (L.nz + 8 + (1 << BRICK_SZ) - 1) >> BRICK_SZ;
    const size_t nbr = (size_t)nbx * nby * nbz;
    auto brick = [&](int o) -> long {
        const int kp = o / L.pxy, r = o - kp * L.pxy, jp = r / L.px, ip = r - jp * L.px;
        return (ip >> BRICK_SX) + nbx * ((jp >> BRICK_SY) + (long)nby * (kp >> BRICK_SZ));
    };
    std::stable_sort(nrm.begin(), nrm.end(), [&](int a, int b) { return brick(a) < brick(b); });
    // solid list (all, and the part inside one halo layer for the normal extrapolation)
    std::vector<int> sperm(ns);
    for (int t = 0; t < ns; ++t) sperm[t] = t;
    {
        std::vector<std::pair<long, long>> skey(ns);
        for (int t = 0; t < ns; ++t) {
            const int o = L.at(s_ijk[3 * t], s_ijk[3 * t + 1], s_ijk[3 * t + 2]);
            skey[t] = {brick(o), T.key(o)};
        }
        std::stable_sort(sperm.begin(), sperm.end(), [&](int a, int b) { return skey[a] < skey[b]; });
    }
    std::vector<int> spos(ns), s1pos;
    std::vector<unsigned> smask_o(ns), s1mask;
    std::vector<double> sw_o(ns), s1w;
    for (int u = 0; u < ns; ++u) {
        const int t = sperm[u];
        const int i = s_ijk[3 * t], j = s_ijk[3 * t + 1], k = s_ijk[3 * t + 2];
        spos[u] = L.at(i, j, k);
        smask_o[u] = smask[t];
        sw_o[u] = s_w[t];
        if (i >= 0 && i <= L.nx + 1 && j >= 0 && j <= L.ny + 1 && k >= 0 && k <= L.nz + 1) {
            s1pos.push_back(spos[u]);
            s1mask.push_back(smask[t]);
            s1w.push_back(s_w[t]);
        }
    }
    std::vector<int> fpos(nf);
    std::vector<double> fd((size_t)nf * 5 + 1, 0.0);
    for (int t = 0; t < nf; ++t) {
        fpos[t] = L.at(f_ijk[3 * t], f_ijk[3 * t + 1], f_ijk[3 * t + 2]);
        for (int m = 0; m < 3; ++m) fd[3 * t + m] = f_nw[3 * t + m];
        // the reference evaluates dcos/dsin of the node's angle on the host CPU; doing it here keeps the
        // device free of a second math library's rounding
        fd[3 * (size_t)nf + t] = std::cos(f_theta[t]);
        fd[4 * (size_t)nf + t] = std::sin(f_theta[t]);
    }
    // fluid boundary nodes are rotated inside the normal kernel: index of each normal-list node in the list
    // (nodes of the list that are solid in the wall array get a zero normal in the reference, so no rotation)
    std::vector<int> wb(nrm.size(), -1);
    {
        std::vector<int> at(L.total, -1);
        for (int t = 0; t < nf; ++t) at[fpos[t]] = t;
        for (size_t t = 0; t < nrm.size(); ++t) wb[t] = at[nrm[t]];
    }
    // the same per node for the fused chain: the fluid boundary index of a non-solid node (as wb), or the
    // extrapolation-list index of a solid boundary node in [0, n+1]^3 as -index - 2
    std::vector<int> role(L.total, -1);
    for (int t = 0; t < nf; ++t)
        if (wp[fpos[t]] == 0) role[fpos[t]] = t;
    for (size_t u = 0; u < s1pos.size(); ++u) role[s1pos[u]] = -(int)u - 2;
    // device buffers
    free_geometry(c);
    // slice q: nc compact slots, then the link slots of direction opp(q), each slot the pair (colour 1, colour
    // 2) of direction q; each slice 256-byte aligned
    const long esz = c->fp32 ? (long)sizeof(float) : (long)sizeof(double);
    long soff[20];
    soff[0] = 0;
    for (int q = 0; q < 19; ++q) {
        const long len = (((long)nc + (q ? nlink[lat::OPP[q]] : 0) + 63) / 64) * 64;
        if (2 * len * esz >= (1L << 32))
            return fail(c, MFLBM_ERR_ARG, "population slice of %ld slots exceeds 32-bit byte offsets", len);
        soff[q + 1] = soff[q] + 2 * len;
    }
    const long fs = (((long)nfl + 63) / 64) * 64;
    const size_t ni = (size_t)nc + nrm.size() + ns + s1pos.size() + nf + 1;
    const size_t nd = (size_t)ns + s1w.size() + fd.size() + 1;
    HIPCHK(c, hipMalloc(&c->compact, esz * soff[19]));
    HIPCHK(c, hipMalloc(&c->force, sizeof(double) * fs * 4));
    HIPCHK(c, hipMalloc(&c->map, 2 * sizeof(int) * L.total));   // map, then pmap
    HIPCHK(c, hipMalloc(&c->ibuf, sizeof(int) * ni));
    HIPCHK(c, hipMalloc(&c->wbuf, sizeof(int) * (wb.size() + 1)));
    if (!wb.empty()) HIPCHK(c, hipMemcpy(c->wbuf, wb.data(), sizeof(int) * wb.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMalloc(&c->mbuf, sizeof(unsigned) * ((size_t)ns + s1mask.size() + 1)));
    HIPCHK(c, hipMalloc(&c->dbuf, sizeof(double) * nd));
    HIPCHK(c, hipMalloc(&c->wmask, sizeof(unsigned long long) * links.wmask.size()));
    HIPCHK(c, hipMalloc(&c->wbase, sizeof(int) * wbase.size()));
    HIPCHK(c, hipMemcpy(c->wmask, links.wmask.data(), sizeof(unsigned long long) * links.wmask.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->wbase, wbase.data(), sizeof(int) * wbase.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemset(c->compact, 0, esz * soff[19]));
    HIPCHK(c, hipMemset(c->force, 0, sizeof(double) * fs * 4));
    // normals at solid nodes are only ever written by the extrapolation; start them from zero so no value
    // of a previous geometry survives (cn_x, cn_y, cn_z, c_norm, curv, u, v, w, rho)
    HIPCHK(c, hipMemset(c->d.cn, 0, sizeof(double) * c->ps * 9));
    if (c->fields32) HIPCHK(c, hipMemset(c->fields32, 0, sizeof(float) * c->ps * 4));
    std::vector<int> hi;
    hi.reserve(ni);
    hi.insert(hi.end(), cpos.begin(), cpos.end());
    hi.insert(hi.end(), nrm.begin(), nrm.end());
    hi.insert(hi.end(), spos.begin(), spos.end());
    hi.insert(hi.end(), s1pos.begin(), s1pos.end());
    hi.insert(hi.end(), fpos.begin(), fpos.end());
    hi.push_back(0);
    std::vector<unsigned> hm(smask_o);
    hm.insert(hm.end(), s1mask.begin(), s1mask.end());
    hm.push_back(0);
    std::vector<double> hd(sw_o);
    hd.insert(hd.end(), s1w.begin(), s1w.end());
    hd.insert(hd.end(), fd.begin(), fd.end());
    hd.push_back(0.0);
    HIPCHK(c, hipMemcpy(c->map, cidx.data(), sizeof(int) * L.total, hipMemcpyHostToDevice));
    {
        std::vector<int> pm(L.total, -1);
        for (int m = 0; m < nfl; ++m) pm[cpos[m]] = m;
        HIPCHK(c, hipMemcpy(c->map + L.total, pm.data(), sizeof(int) * L.total, hipMemcpyHostToDevice));
    }
    HIPCHK(c, hipMalloc(&c->role, sizeof(int) * L.total));
    HIPCHK(c, hipMemcpy(c->role, role.data(), sizeof(int) * L.total, hipMemcpyHostToDevice));
    // brick flags (Dev): notm and notp (two buffers each), bstatic, nzb, nzd.  The band flags, nzb and nzd start
    // set, so the first chain and lattice update evaluate everything; bstatic from the
    // walls: bricks with a non-solid node outside the interior, and bricks with an interior node within 4 of a
    // periodic face
    std::vector<unsigned char> hstatic(nbr, 0);
    {
        const int py = L.ny + 8, pz = L.nz + 8;
        for (int kp = 0; kp < pz; ++kp)
            for (int jp = 0; jp < py; ++jp)
                for (int ip = 0; ip < L.px; ++ip) {
                    const int i = ip - XO, j = jp - YO, k = kp - ZO;
                    const bool inside = i >= 1 && i <= L.nx && j >= 1 && j <= L.ny && k >= 1 && k <= L.nz;
                    const long o = (long)ip + (long)jp * L.px + (long)kp * L.pxy;
                    const bool edge = inside && ((c->p.kper == 1 && (k <= 4 || k > L.nz - 4)) ||
                                                 (c->p.jper == 1 && (j <= 4 || j > L.ny - 4)));
                    if ((!inside && wp[o] == 0) || edge)
                        hstatic[(ip >> BRICK_SX) + nbx * ((jp >> BRICK_SY) + (size_t)nby * (kp >> BRICK_SZ))] = 1;
                }
    }
    // chain lists: the brick offsets of the sorted lists, which of them a brick has entries in (hasb), and the
    // work lists with their counters (zero: the first plan fills set 0)
    const int ntiles = ((nbx + TBX - 1) / TBX) * ((nby + TBY - 1) / TBY) * ((nbz + TBZ - 1) / TBZ);
    const int lcap = (ntiles + NXCD - 1) / NXCD * BLOCK;   // items per sub-list: the bricks of its tiles
    const size_t lw = (size_t)NXCD * lcap;
    std::vector<int> hc(4 * NWORK * lw + 3 * (nbr + 1) + 2 * NWORK * NXCD, 0);
    std::vector<unsigned char> hasb(nbr, 0);
    {
        const std::vector<int>* lists[3] = {&nrm, &spos, &s1pos};
        for (int l = 0; l < 3; ++l) {
            int* off = hc.data() + 4 * NWORK * lw + (size_t)l * (nbr + 1);
            for (int o : *lists[l]) ++off[brick(o) + 1];
            for (size_t b = 0; b < nbr; ++b) {
                if (off[b + 1] > 64) return fail(c, MFLBM_ERR_ARG, "brick with more than 64 chain entries");
                if (off[b + 1]) hasb[b] |= (unsigned char)(1 << l);
                off[b + 1] += off[b];
            }
        }
    }
    HIPCHK(c, hipMalloc(&c->qbuf, 8 * nbr));
    HIPCHK(c, hipMemset(c->qbuf, 1, 7 * nbr));
    HIPCHK(c, hipMemcpy(c->qbuf + 4 * nbr, hstatic.data(), nbr, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->qbuf + 7 * nbr, hasb.data(), nbr, hipMemcpyHostToDevice));
    HIPCHK(c, hipMalloc(&c->cbuf, sizeof(int) * hc.size()));
    HIPCHK(c, hipMemcpy(c->cbuf, hc.data(), sizeof(int) * hc.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMalloc(&c->phic, sizeof(double) * ((size_t)nfl + 1)));
    HIPCHK(c, hipMemset(c->phic, 0, sizeof(double) * ((size_t)nfl + 1)));
    c->phi_stale = false;
    c->nbricks = (long)nbr;
    c->lcap = lcap;
    c->band_cur = 0;
    c->full_chain = true;
    c->solid_zeroed = false;
    HIPCHK(c, hipMemcpy(c->ibuf, hi.data(), sizeof(int) * hi.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->mbuf, hm.data(), sizeof(unsigned) * hm.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->dbuf, hd.data(), sizeof(double) * hd.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->walls, wp.data(), (size_t)L.total, hipMemcpyHostToDevice));
    if (w_in) HIPCHK(c, hipMemcpy(c->d.w_in, w_in, sizeof(double) * (L.nx + 2) * (L.ny + 2), hipMemcpyHostToDevice));
    Dev& d = c->d;
    // FP32 storage keeps float slices behind the same pointers (the kernels templated on the storage type cast)
    // p[q] colour 1 and p[19 + q] colour 2 of direction q: the two halves of the pairs (stride two elements)
    for (int q = 0; q < 19; ++q) {
        d.S.p[q] = (double*)((char*)c->compact + esz * soff[q]);
        d.S.p[19 + q] = (double*)((char*)c->compact + esz * (soff[q] + 1));
    }
    d.map = c->map;
    d.pmap = c->map + L.total;
    d.phic = c->phic;
    d.cpos = c->ibuf;
    d.nf = nfl;
    d.nc = nc;
    d.force = c->force;
    d.fs = fs;
    d.wmask = c->wmask;
    d.wbase = c->wbase;
    d.notm = c->qbuf;
    d.notp = c->qbuf + 2 * nbr;
    d.bstatic = c->qbuf + 4 * nbr;
    d.nzb = c->qbuf + 5 * nbr;
    d.nzd = c->qbuf + 6 * nbr;
    d.nbx = nbx;
    d.nby = nby;
    d.nbz = nbz;
    const int* ip = c->ibuf + nc;
    c->nrm = NodeList{(int)nrm.size(), ip};
    ip += nrm.size();
    c->sl = SolidList{ns, ip, c->mbuf, c->dbuf};
    ip += ns;
    c->sl1 = SolidList{(int)s1pos.size(), ip, c->mbuf + ns, c->dbuf + ns};
    ip += s1pos.size();
    const double* dp = c->dbuf + ns + s1w.size();
    c->fl = FluidList{nf, ip, dp, dp + 3 * (size_t)nf, dp + 4 * (size_t)nf};
    HIPCHK(c, hipDeviceSynchronize());
    c->have_geom = true;
    return MFLBM_OK;
}

extern "C" int mflbm_upload_state(mflbm_ctx* c, const double* const* f, const double* const* g, const double* phi,
                                  const double* f_convec, const double* g_convec, const double* phi_convec) {
    if (!c) return MFLBM_ERR_ARG;
    if (!c->have_geom) return fail(c, MFLBM_ERR_STATE, "upload the geometry before the state");
    if (!f || !g || !phi) return fail(c, MFLBM_ERR_ARG, "missing state array");
    for (int q = 0; q < 19; ++q)
        if (!f[q] || !g[q]) return fail(c, MFLBM_ERR_ARG, "population pointer %d is NULL", q);
    const Layout& L = c->L;
    std::vector<double> buf(L.total, 0.0);
    for (int q = 0; q < 38; ++q) {
        pack_host(L, 1, q < 19 ? f[q] : g[q - 19], buf.data());
        HIPCHK(c, hipMemcpy(c->dense_pdf + q * c->ps, buf.data(), sizeof(double) * L.total, hipMemcpyHostToDevice));
    }
    std::fill(buf.begin(), buf.end(), 0.0);
    pack_host(L, 4, phi, buf.data());
    HIPCHK(c, hipMemcpy(c->d.phi, buf.data(), sizeof(double) * L.total, hipMemcpyHostToDevice));
    const size_t plane = (size_t)(L.nx + 2) * (L.ny + 2);
    if (f_convec) HIPCHK(c, hipMemcpy(c->d.cf, f_convec, sizeof(double) * plane * 19, hipMemcpyHostToDevice));
    if (g_convec) HIPCHK(c, hipMemcpy(c->d.cg, g_convec, sizeof(double) * plane * 19, hipMemcpyHostToDevice));
    if (phi_convec) HIPCHK(c, hipMemcpy(c->d.cphi, phi_convec, sizeof(double) * plane, hipMemcpyHostToDevice));
    HIPCHK(c, hipDeviceSynchronize());
    {
        const int rs = c->fp32 ? gather_state<float>(c) : gather_state<double>(c);
        if (rs) return rs;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // phi_old reads as zero until a monitor call updates it (a restarted run's first phase-field monitor then
    // sees the whole field as changed, as the reference's unset phi_old does)
    HIPCHK(c, hipMemset(c->phi_old, 0, sizeof(double) * c->ps));
    c->have_state = true;
    c->full_chain = true;
    return MFLBM_OK;
}

extern "C" int mflbm_init_state_device(mflbm_ctx* c, int32_t option, double interface_z0) {
    if (!c) return MFLBM_ERR_ARG;
    if (!c->have_geom) return fail(c, MFLBM_ERR_STATE, "upload the geometry before the state");
    if (option < 1 || option > 5)
        return fail(c, MFLBM_ERR_ARG, "initial_fluid_distribution_option %d not supported (1..5)", option);
    const Layout& L = c->L;
    const mflbm_params& p = c->p;
    const bool open_z = p.kper == 0 && p.wall_z_min == 0 && p.wall_z_max == 0;
    HIPCHK(c, hipMemsetAsync(c->d.phi, 0, sizeof(double) * L.total, c->stream));
    const long nb = (long)(L.nx + 8) * (L.ny + 8) * (L.nz + 8);
    LAUNCH(c, k_init_dense, nblocks(nb), c->d, c->dense_pdf, c->ps, (int)option, interface_z0, p.phi_inlet,
           open_z ? 1 : 0);
    if (p.outlet_bc == 1) LAUNCH(c, k_init_convec, nblocks((long)(L.nx + 2) * (L.ny + 2)), c->d, (const double*)c->dense_pdf, c->ps);
    {
        const int rs = c->fp32 ? gather_state<float>(c) : gather_state<double>(c);
        if (rs) return rs;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // phi_old reads as zero until a monitor call updates it (a restarted run's first phase-field monitor then
    // sees the whole field as changed, as the reference's unset phi_old does)
    HIPCHK(c, hipMemset(c->phi_old, 0, sizeof(double) * c->ps));
    c->have_state = true;
    c->full_chain = true;
    return MFLBM_OK;
}

// phi at the solid nodes as the reference holds it: the chain leaves the extrapolated phi of skipped bricks as it
// was (no evaluated normal reads it), so before phi leaves the device every entry is extrapolated from the
// current fluid phi, as the last chain of the reference did, and zeroed again inside solids when the macroscopic
// pass has run since that chain (compute_macro_vars, Misc.F90:423)
int refresh_solid_phi(mflbm_ctx* c) {
    if (!c->have_geom || !c->have_state) return MFLBM_OK;
    const int re = ensure_dense_phi(c);
    if (re) return re;
    LAUNCH(c, k_phi_extrapolate, NXCD * xcd_per(c->sl.n), c->d, c->sl, xcd_per(c->sl.n));
    if (c->solid_zeroed) LAUNCH(c, k_zero_solid_phi, nblocks((long)c->L.nx * c->L.ny * c->L.nz), c->d);
    return MFLBM_OK;
}

extern "C" int mflbm_download_state(mflbm_ctx* c, double* const* f, double* const* g, double* phi, double* f_convec,
                                    double* g_convec, double* phi_convec) {
    if (!c) return MFLBM_ERR_ARG;
    if (!c->have_state) return fail(c, MFLBM_ERR_STATE, "no state uploaded");
    const Layout& L = c->L;
    const bool folded = c->fold.y || c->fold.z;
    if (!folded || c->stepped) {   // (before the first step of a folded domain the staging holds the state as given)
        const int rs = c->fp32 ? scatter_state<float>(c) : scatter_state<double>(c);
        if (rs) return rs;
        if (folded)
            LAUNCH(c, k_fold_dense, fold_blocks(L), L, c->fold, (const int*)c->map, c->d.nf, c->dense_pdf, c->ps, 2);
    }
    if (phi) {
        const int rr = refresh_solid_phi(c);
        if (rr) return rr;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<double> buf(L.total);
    for (int q = 0; q < 38; ++q) {
        double* dst = q < 19 ? (f ? f[q] : nullptr) : (g ? g[q - 19] : nullptr);
        if (!dst) continue;
        HIPCHK(c, hipMemcpy(buf.data(), c->dense_pdf + q * c->ps, sizeof(double) * L.total, hipMemcpyDeviceToHost));
        unpack_host(L, 1, buf.data(), dst);
    }
    if (phi) {
        HIPCHK(c, hipMemcpy(buf.data(), c->d.phi, sizeof(double) * L.total, hipMemcpyDeviceToHost));
        unpack_host(L, 4, buf.data(), phi);
    }
    const size_t plane = (size_t)(L.nx + 2) * (L.ny + 2);
    if (f_convec) HIPCHK(c, hipMemcpy(f_convec, c->d.cf, sizeof(double) * plane * 19, hipMemcpyDeviceToHost));
    if (g_convec) HIPCHK(c, hipMemcpy(g_convec, c->d.cg, sizeof(double) * plane * 19, hipMemcpyDeviceToHost));
    if (phi_convec) HIPCHK(c, hipMemcpy(phi_convec, c->d.cphi, sizeof(double) * plane, hipMemcpyDeviceToHost));
    return MFLBM_OK;
}

extern "C" int mflbm_color_gradient(mflbm_ctx* c) {
    if (!c) return MFLBM_ERR_ARG;
    if (!c->have_geom || !c->have_state) return fail(c, MFLBM_ERR_STATE, "upload geometry and state first");
    c->full_chain = true;
    return launch_gradient(c);
}

extern "C" int mflbm_step(mflbm_ctx* c, int32_t ntime) {
    if (!c) return MFLBM_ERR_ARG;
    if (!c->have_geom || !c->have_state) return fail(c, MFLBM_ERR_STATE, "upload geometry and state first");
    return launch_step(c, ntime);
}

extern "C" int mflbm_run(mflbm_ctx* c, int32_t ntime0, int32_t nsteps) {
    if (!c) return MFLBM_ERR_ARG;
    if (!c->have_geom || !c->have_state) return fail(c, MFLBM_ERR_STATE, "upload geometry and state first");
    if (nsteps < 0) return fail(c, MFLBM_ERR_ARG, "nsteps < 0");
    for (int s = 0; s < nsteps; ++s) {
        int rc = launch_step(c, ntime0 + s);
        if (rc) return rc;
    }
    return MFLBM_OK;
}

extern "C" int mflbm_compute_macro(mflbm_ctx* c) {
    if (!c) return MFLBM_ERR_ARG;
    if (!c->have_state) return fail(c, MFLBM_ERR_STATE, "no state uploaded");
    const int re = ensure_dense_phi(c);   // the monitor and the downloads read the dense phi
    if (re) return re;
    // the force input of the macroscopic velocity comes from the current normal field
    // (the fused chain keeps the force arrays current; the list chain leaves only the dense normals)
    if (!c->chain_fused) {
        if (c->fp32) LAUNCH(c, k_curvature<float>, NXCD * xcd_per(c->d.nf), c->d, c->p.gamma, xcd_per(c->d.nf));
        else LAUNCH(c, k_curvature<double>, NXCD * xcd_per(c->d.nf), c->d, c->p.gamma, xcd_per(c->d.nf));
    }
    if (c->fp32) LAUNCH(c, k_macro<float>, nblocks(c->d.nf), c->d, c->p.force_z, c->u, c->v, c->w, c->rho);
    else LAUNCH(c, k_macro<double>, nblocks(c->d.nf), c->d, c->p.force_z, c->u, c->v, c->w, c->rho);
    LAUNCH(c, k_zero_solid_phi, nblocks((long)c->L.nx * c->L.ny * c->L.nz), c->d);
    c->solid_zeroed = true;
    return MFLBM_OK;
}

extern "C" int mflbm_monitor(mflbm_ctx* c, double* layers, mflbm_monitor_sums* sums, int32_t update_phi_old) {
    if (!c) return MFLBM_ERR_ARG;
    if (!c->have_state) return fail(c, MFLBM_ERR_STATE, "no state uploaded");
    int rc = mflbm_compute_macro(c);
    if (rc) return rc;
    const Layout& L = c->L;
    double* lay = c->red;
    double* part = c->red + 7 * (size_t)L.nz;
    LAUNCH(c, k_monitor_layers, (unsigned)L.nz, c->d, c->u, c->v, c->w, c->rho, c->phi_old,
           lay, part);
    std::vector<double> h((size_t)L.nz * (7 + NPART + 3 - 7));
    HIPCHK(c, hipMemcpyAsync(h.data(), c->red, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (layers) std::memcpy(layers, h.data(), sizeof(double) * 7 * L.nz);
    if (sums) {
        const double* P = h.data() + 7 * (size_t)L.nz;
        auto row = [&](int m) { return P + (size_t)m * L.nz; };
        mflbm_monitor_sums s{};
        // layer partials combined in layer order on the host (deterministic)
        for (int k = 0; k < L.nz; ++k) {
            s.usq1 += row(0)[k];
            s.usq2 += row(1)[k];
            s.pre_w += row(2)[k];
            s.n_w += (int64_t)row(3)[k];
            s.pre_nw += row(4)[k];
            s.umax2 = k == 0 ? row(5)[k] : (row(5)[k] > s.umax2 ? row(5)[k] : s.umax2);
            s.dphi_max = k == 0 ? row(6)[k] : (row(6)[k] > s.dphi_max ? row(6)[k] : s.dphi_max);
            s.n_nw += (int64_t)row(7)[k];
            s.vol1_full += h[2 * (size_t)L.nz + k];
            s.vol2_full += h[3 * (size_t)L.nz + k];
        }
        // breakthrough count at z = nz-1 (Monitor.F90:472-507)
        std::vector<double> phil((size_t)L.nx * L.ny);
        std::vector<int8_t> wl((size_t)L.nx * L.ny);
        for (int j = 1; j <= L.ny; ++j) {
            HIPCHK(c, hipMemcpy(phil.data() + (size_t)(j - 1) * L.nx, c->d.phi + L.at(1, j, L.nz - 1), sizeof(double) * L.nx,
                                hipMemcpyDeviceToHost));
            HIPCHK(c, hipMemcpy(wl.data() + (size_t)(j - 1) * L.nx, c->walls + L.at(1, j, L.nz - 1), (size_t)L.nx,
                                hipMemcpyDeviceToHost));
        }
        for (size_t m = 0; m < phil.size(); ++m)
            if (wl[m] == 0 && phil[m] > 0.0) s.outlet_phase1++;
        *sums = s;
    }
    if (update_phi_old) {
        LAUNCH(c, k_copy_phi_old, nblocks(L.total), c->d, c->phi_old);
    }
    return MFLBM_OK;
}

extern "C" int mflbm_download_macro(mflbm_ctx* c, double* u, double* v, double* w, double* rho, double* phi) {
    if (!c) return MFLBM_ERR_ARG;
    const Layout& L = c->L;
    if (phi) {
        const int rr = refresh_solid_phi(c);
        if (rr) return rr;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<double> buf(L.total);
    double* src[5] = {c->u, c->v, c->w, c->rho, c->d.phi};
    double* dst[5] = {u, v, w, rho, phi};
    for (int m = 0; m < 5; ++m) {
        if (!dst[m]) continue;
        HIPCHK(c, hipMemcpy(buf.data(), src[m], sizeof(double) * L.total, hipMemcpyDeviceToHost));
        unpack_host(L, m == 4 ? 4 : 1, buf.data(), dst[m]);
    }
    return MFLBM_OK;
}

extern "C" int mflbm_download_gradient(mflbm_ctx* c, double* cn_x, double* cn_y, double* cn_z, double* c_norm, double* curv) {
    if (!c) return MFLBM_ERR_ARG;
    const Layout& L = c->L;
    // the step keeps curvature only at fluid nodes (folded into the force input); the dense field the
    // reference holds is evaluated from the current normals on request
    if (c->chain_fused) {   // the fused chain does not keep the dense normal field: evaluate it now, over the
                            // lists of the last chain (every brick: the fused chain plans a full chain)
        const int re = ensure_dense_phi(c);
        if (re) return re;
        const int set = c->band_cur ^ 1;
        const int rc = c->fp32 ? launch_normals_listed<float>(c, set) : launch_normals_listed<double>(c, set);
        if (rc) return rc;
    }
    if (curv) {
        if (c->fp32) LAUNCH(c, k_curvature_dense<float>, nblocks((long)L.nx * L.ny * L.nz), c->d);
        else LAUNCH(c, k_curvature_dense<double>, nblocks((long)L.nx * L.ny * L.nz), c->d);
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<double> buf(L.total);
    double* dst[5] = {cn_x, cn_y, cn_z, c_norm, curv};
    if (cn_x || cn_y || cn_z || c_norm) {   // the records, split into the components (FP32 widened)
        std::vector<double> rec(4 * (size_t)L.total);
        if (c->fp32) {
            std::vector<float> rec32(4 * (size_t)L.total);
            HIPCHK(c, hipMemcpy(rec32.data(), c->d.cn32, sizeof(float) * rec32.size(), hipMemcpyDeviceToHost));
            for (size_t k = 0; k < rec.size(); ++k) rec[k] = (double)rec32[k];
        } else {
            HIPCHK(c, hipMemcpy(rec.data(), c->d.cn, sizeof(double) * rec.size(), hipMemcpyDeviceToHost));
        }
        for (int m = 0; m < 4; ++m) {
            if (!dst[m]) continue;
            for (long k = 0; k < L.total; ++k) buf[k] = rec[4 * (size_t)k + m];
            unpack_host(L, 2, buf.data(), dst[m]);
        }
    }
    if (curv) {
        HIPCHK(c, hipMemcpy(buf.data(), c->d.curv, sizeof(double) * L.total, hipMemcpyDeviceToHost));
        unpack_host(L, 1, buf.data(), curv);
    }
    return MFLBM_OK;
}

extern "C" int mflbm_synchronize(mflbm_ctx* c) {
    if (!c) return MFLBM_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MFLBM_OK;
}

namespace {
int fold_events(mflbm_ctx* c) {
    if (c->ev_used == 0) return MFLBM_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (size_t b = 0; b + 4 <= c->ev_used; b += 4) {
        float a = 0, g = 0, s = 0;
        HIPCHK(c, hipEventElapsedTime(&a, c->ev[b], c->ev[b + 1]));
        HIPCHK(c, hipEventElapsedTime(&g, c->ev[b + 2], c->ev[b + 3]));
        HIPCHK(c, hipEventElapsedTime(&s, c->ev[b], c->ev[b + 3]));
        c->t_sum[0] += a;
        c->t_sum[1] += g;
        c->t_sum[2] += s;
        c->t_cnt[0]++;
        c->t_cnt[1]++;
        c->t_cnt[2]++;
    }
    c->ev_used = 0;
    return MFLBM_OK;
}
}  // namespace

extern "C" int mflbm_timing_enable(mflbm_ctx* c, int32_t enable) {
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

extern "C" int mflbm_timing(mflbm_ctx* c, int32_t which, double* mean_ms, int64_t* count) {
    if (!c || which < 0 || which > 2) return MFLBM_ERR_ARG;
    int rc = fold_events(c);
    if (rc) return rc;
    if (mean_ms) *mean_ms = c->t_cnt[which] ? c->t_sum[which] / c->t_cnt[which] : 0.0;
    if (count) *count = c->t_cnt[which];
    return MFLBM_OK;
}
