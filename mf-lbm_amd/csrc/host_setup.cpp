// Host-side set-up of a run: local wall array, boundary-node lists with wall normals, initial fields and
// the inlet velocity profile.  These run once per run on the host and feed the device state; their
// arithmetic follows the reference expression by expression so the device path starts from the same bits.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

#include "geometry.h"
#include "lattice.h"
#include "mflbm.h"

namespace {

// (i, j, k) -> offset in an array with halo h in every direction, x fastest.
struct Grid {
    int nx, ny, nz, h;
    long sx, sy;
    Grid(int nx_, int ny_, int nz_, int h_) : nx(nx_), ny(ny_), nz(nz_), h(h_) {
        sx = nx + 2 * h;
        sy = (long)(ny + 2 * h) * sx;
    }
    long at(int i, int j, int k) const { return (long)(i + h - 1) + (long)(j + h - 1) * sx + (long)(k + h - 1) * sy; }
    long size() const { return sy * (nz + 2 * h); }
};

bool bad_params(const mflbm_params* p) {
    return !p || p->nx < 3 || p->ny < 3 || p->nz < 3;
}

}  // namespace

extern "C" int mflbm_setup_walls(const mflbm_params* p, const int8_t* wg, int8_t* walls_h2) {
    // set_walls for a single rank (Misc.F90:24-36, 130-210) followed by the periodic transport of the
    // wall halo (ztransport_walls(0,0,2), ytransport_walls(0,2,2), xtransport_walls(2,2,2); x is never
    // periodic so the x transport changes nothing).
    if (bad_params(p) || !wg || !walls_h2) return MFLBM_ERR_ARG;
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    Grid G(nx, ny, nz, 2);
    std::memset(walls_h2, 0, (size_t)G.size());
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                walls_h2[G.at(i, j, k)] = wg[(long)(i - 1) + (long)(j - 1) * nx + (long)(k - 1) * nx * ny];
    for (int k = -1; k <= nz + 2; ++k)
        for (int j = -1; j <= ny + 2; ++j)
            for (int i = -1; i <= nx + 2; ++i) {
                int8_t& w = walls_h2[G.at(i, j, k)];
                if (k <= 1 && p->wall_z_min == 1) w = 1;
                if (k >= nz && p->wall_z_max == 1) w = 1;
                if (i <= 1 && p->wall_x_min == 1) w = 1;
                if (i >= nx && p->wall_x_max == 1) w = 1;
                if (j <= 1 && p->wall_y_min == 1) w = 1;
                if (j >= ny && p->wall_y_max == 1) w = 1;
            }
    if (p->kper == 1) {  // z halo of the interior x/y range from the periodic image
        for (int l = 1; l <= 2; ++l)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    walls_h2[G.at(i, j, l - 2)] = walls_h2[G.at(i, j, nz + l - 2)];
                    walls_h2[G.at(i, j, nz + l)] = walls_h2[G.at(i, j, l)];
                }
    }
    if (p->jper == 1) {  // y halo for k over the full z range including its halo
        for (int k = -1; k <= nz + 2; ++k)
            for (int l = 1; l <= 2; ++l)
                for (int i = 1; i <= nx; ++i) {
                    walls_h2[G.at(i, l - 2, k)] = walls_h2[G.at(i, ny + l - 2, k)];
                    walls_h2[G.at(i, ny + l, k)] = walls_h2[G.at(i, l, k)];
                }
    }
    return MFLBM_OK;
}

extern "C" int mflbm_geometry_new(const mflbm_params* p, const int8_t* wg, double theta, mflbm_geometry** out) {
    // geometry_preprocessing_new (Geometry_preprocessing.F90:9-512), single rank: extended wall copy with
    // 10 ghost layers, 1/2 (solid) and 0/-1 (fluid) boundary marks, four passes of 27-point smoothing, then
    // the solid list (with fluid neighbours and weights) and the fluid list (ISO8 normals) over the range
    // that covers the phi halo, finally restricted to 3 (solid) / 2 (fluid) ghost layers.
    if (bad_params(p) || !wg || !out) return MFLBM_ERR_ARG;
    *out = nullptr;
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    const int GL = geo::GL;
    if ((p->kper == 1 && nz < GL) || (p->jper == 1 && ny < GL)) return MFLBM_ERR_ARG;
    Grid G(nx, ny, nz, GL);
    std::vector<int8_t> t(G.size(), 0);
    std::vector<double> s1(G.size()), s2(G.size());
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                t[G.at(i, j, k)] = wg[(long)(i - 1) + (long)(j - 1) * nx + (long)(k - 1) * nx * ny];
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i)
            for (int l = 1 - GL; l <= 0; ++l) {
                if (p->kper == 0) {
                    t[G.at(i, j, l)] = t[G.at(i, j, 1)];
                    t[G.at(i, j, nz + GL + l)] = t[G.at(i, j, nz)];
                } else {
                    t[G.at(i, j, l)] = t[G.at(i, j, nz + l)];
                    t[G.at(i, j, nz + GL + l)] = t[G.at(i, j, GL + l)];
                }
            }
    for (int k = 1 - GL; k <= nz + GL; ++k)
        for (int i = 1; i <= nx; ++i)
            for (int l = 1 - GL; l <= 0; ++l) {
                if (p->jper == 0) {
                    t[G.at(i, l, k)] = t[G.at(i, 1, k)];
                    t[G.at(i, ny + GL + l, k)] = t[G.at(i, ny, k)];
                } else {
                    t[G.at(i, l, k)] = t[G.at(i, ny + l, k)];
                    t[G.at(i, ny + GL + l, k)] = t[G.at(i, GL + l, k)];
                }
            }
    for (int k = 1 - GL; k <= nz + GL; ++k)
        for (int j = 1 - GL; j <= ny + GL; ++j)
            for (int l = 1 - GL; l <= 0; ++l) {  // x is never periodic
                t[G.at(l, j, k)] = t[G.at(1, j, k)];
                t[G.at(nx + GL + l, j, k)] = t[G.at(nx, j, k)];
            }
    for (long n = 0; n < G.size(); ++n) s1[n] = s2[n] = (double)t[n];
    // boundary marks; the fluid/solid class of every node is unchanged by the marks, so the in-place sweep
    // gives the same result in any order
    for (int k = 2 - GL; k <= nz + GL - 1; ++k)
        for (int j = 2 - GL; j <= ny + GL - 1; ++j)
            for (int i = 2 - GL; i <= nx + GL - 1; ++i) {
                int8_t& c = t[G.at(i, j, k)];
                if (c == 1) {
                    for (int q = 1; q <= 18; ++q)
                        if (t[G.at(i + lat::EX[q], j + lat::EY[q], k + lat::EZ[q])] <= 0) { c = 2; break; }
                }
                if (c == 0) {
                    for (int q = 1; q <= 18; ++q)
                        if (t[G.at(i + lat::EX[q], j + lat::EY[q], k + lat::EZ[q])] >= 1) { c = -1; break; }
                }
            }
    for (int it = 0; it < 4; ++it) {
        auto S1 = [&](int i, int j, int k) { return s1[G.at(i, j, k)]; };
        for (int k = 2 - GL; k <= nz + GL - 1; ++k)
            for (int j = 2 - GL; j <= ny + GL - 1; ++j)
                for (int i = 2 - GL; i <= nx + GL - 1; ++i) s2[G.at(i, j, k)] = geo::smooth27(S1, i, j, k);
        for (int k = 2 - GL; k <= nz + GL - 1; ++k)
            for (int j = 2 - GL; j <= ny + GL - 1; ++j)
                for (int i = 2 - GL; i <= nx + GL - 1; ++i) s1[G.at(i, j, k)] = s2[G.at(i, j, k)];
    }
    mflbm_geometry* gm = new (std::nothrow) mflbm_geometry();
    if (!gm) return MFLBM_ERR_NOMEM;
    const double* W = s2.data();
    auto S = [&](int i, int j, int k) { return W[G.at(i, j, k)]; };
    for (int k = -3; k <= nz + 4; ++k)
        for (int j = -3; j <= ny + 4; ++j)
            for (int i = -3; i <= nx + 4; ++i) {
                const int8_t c = t[G.at(i, j, k)];
                if (c == 2) {
                    if (i < -2 || i > nx + 3 || j < -2 || j > ny + 3 || k < -2 || k > nz + 3) continue;
                    int nbr[18] = {0};
                    int cnt = 0;
                    double w = 0.0;
                    for (int q = 1; q <= 18; ++q)
                        if (t[G.at(i + lat::EX[q], j + lat::EY[q], k + lat::EZ[q])] <= 0) {
                            w = w + lat::W[q];
                            nbr[cnt++] = q;
                        }
                    gm->s_ijk.push_back(i);
                    gm->s_ijk.push_back(j);
                    gm->s_ijk.push_back(k);
                    gm->s_nfl.push_back(cnt);
                    for (int q = 0; q < 18; ++q) gm->s_nbr.push_back(nbr[q]);
                    gm->s_w.push_back(w);
                } else if (c == -1) {
                    if (i < -1 || i > nx + 2 || j < -1 || j > ny + 2 || k < -1 || k > nz + 2) continue;
                    double nwx, nwy, nwz;
                    geo::iso8_gradient(S, i, j, k, nwx, nwy, nwz);
                    double tmp = 1.0 / (std::sqrt(nwx * nwx + nwy * nwy + nwz * nwz) + lat::EPS);
                    gm->f_ijk.push_back(i);
                    gm->f_ijk.push_back(j);
                    gm->f_ijk.push_back(k);
                    gm->f_nw.push_back(nwx * tmp);
                    gm->f_nw.push_back(nwy * tmp);
                    gm->f_nw.push_back(nwz * tmp);
                    gm->f_theta.push_back(theta);
                }
            }
    *out = gm;
    return MFLBM_OK;
}

extern "C" int mflbm_geometry_counts(const mflbm_geometry* g, int32_t* ns, int32_t* nf) {
    if (!g || !ns || !nf) return MFLBM_ERR_ARG;
    *ns = (int32_t)g->s_nfl.size();
    *nf = (int32_t)g->f_theta.size();
    return MFLBM_OK;
}

extern "C" int mflbm_geometry_export(const mflbm_geometry* g, int32_t* s_ijk, int32_t* s_nfl, int32_t* s_nbr,
                                     double* s_w, int32_t* f_ijk, double* f_nw, double* f_theta) {
    if (!g) return MFLBM_ERR_ARG;
    auto cp = [](auto* dst, const auto& src) {
        if (dst && !src.empty()) std::memcpy(dst, src.data(), src.size() * sizeof(src[0]));
    };
    cp(s_ijk, g->s_ijk);
    cp(s_nfl, g->s_nfl);
    cp(s_nbr, g->s_nbr);
    cp(s_w, g->s_w);
    cp(f_ijk, g->f_ijk);
    cp(f_nw, g->f_nw);
    cp(f_theta, g->f_theta);
    return MFLBM_OK;
}

extern "C" void mflbm_geometry_free(mflbm_geometry* g) { delete g; }

extern "C" int mflbm_init_fields(const mflbm_params* p, int32_t option, double z0, const double* u0,
                                 const double* v0, const double* w0, double* const* f, double* const* g,
                                 double* phi, double* f_convec, double* g_convec, double* phi_convec) {
    // initialization_new_multi (Init_multiphase.F90:243-354) and initialization_new_multi_pdf (:357-470)
    if (bad_params(p) || !f || !g || !phi) return MFLBM_ERR_ARG;
    if (option < 1 || option > 5) return MFLBM_ERR_ARG;  // option 6 draws unseeded random numbers
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    Grid P(nx, ny, nz, 4), F(nx, ny, nz, 1);
    const bool open_z = p->kper == 0 && p->wall_z_min == 0 && p->wall_z_max == 0;
    for (int k = -3; k <= nz + 4; ++k)
        for (int j = -3; j <= ny + 4; ++j)
            for (int i = -3; i <= nx + 4; ++i) {
                const double x = i, y = j, z = k;
                double v;
                if (option == 1) {
                    v = -1.0;
                    if (z <= z0) v = 1.0;
                } else if (option == 2) {
                    v = 1.0;
                    if (z <= z0) v = -1.0;
                } else if (option == 3 || option == 4) {
                    const double dx = x - (nx + 1) * 0.0, dz = z - (nz + 1) * 0.5, dy = y - (ny + 1) * 0.5;
                    const bool in = dx * dx + dz * dz + dy * dy <= z0 * z0;
                    v = option == 3 ? (in ? 1.0 : -1.0) : (in ? -1.0 : 1.0);
                } else {
                    const double dx = x - (nx + 1) * 0.5, dz = z - (nz + 1) * 0.5, dy = y - (ny + 1) * 0.5;
                    v = dx * dx + dz * dz + dy * dy <= z0 * z0 ? 1.0 : -1.0;
                }
                if (open_z && z <= 0) v = p->phi_inlet;
                phi[P.at(i, j, k)] = v;
            }
    const double W0 = lat::W[0], W1 = lat::W[1], W2 = lat::W[7];
    // per colour and direction over contiguous arrays; the expression of each population is the reference's
    // with the node's velocity (zero for a run started at rest)
    const long npdf = F.size();
    std::vector<double> rho1(npdf), rho2(npdf), usq(npdf);
    for (int k = 0; k <= nz + 1; ++k)
        for (int j = 0; j <= ny + 1; ++j)
            for (int i = 0; i <= nx + 1; ++i) {
                const long n = F.at(i, j, k);
                const double u = u0 ? u0[n] : 0.0, v = v0 ? v0[n] : 0.0, w = w0 ? w0[n] : 0.0;
                const double rho = 1.0;
                const double ph = phi[P.at(i, j, k)];
                usq[n] = u * u + v * v + w * w;
                rho1[n] = rho * (1.0 + ph) * 0.5;
                rho2[n] = rho * (1.0 - ph) * 0.5;
            }
    for (int c = 0; c < 2; ++c) {
        double* const* d = c == 0 ? f : g;
        const double* rr = c == 0 ? rho1.data() : rho2.data();
        for (int q = 0; q < 19; ++q) {
            double* out = d[q];
            const double wq = q == 0 ? W0 : (q <= 6 ? W1 : W2);
            for (long n = 0; n < npdf; ++n) {
                const double u = u0 ? u0[n] : 0.0, v = v0 ? v0[n] : 0.0, w = w0 ? w0[n] : 0.0;
                double eu, uu;   // e_q . u, and the term squared by the reference (u_x for q = 1, 2, ...)
                switch (q) {
                    case 0: eu = 0.0; uu = 0.0; break;
                    case 1: eu = u; uu = u; break;
                    case 2: eu = -u; uu = u; break;
                    case 3: eu = v; uu = v; break;
                    case 4: eu = -v; uu = v; break;
                    case 5: eu = w; uu = w; break;
                    case 6: eu = -w; uu = w; break;
                    case 7: eu = u + v; uu = eu; break;
                    case 8: eu = -u + v; uu = eu; break;
                    case 9: eu = u - v; uu = eu; break;
                    case 10: eu = -u - v; uu = eu; break;
                    case 11: eu = u + w; uu = eu; break;
                    case 12: eu = -u + w; uu = eu; break;
                    case 13: eu = u - w; uu = eu; break;
                    case 14: eu = -u - w; uu = eu; break;
                    case 15: eu = v + w; uu = eu; break;
                    case 16: eu = -v + w; uu = eu; break;
                    case 17: eu = v - w; uu = eu; break;
                    default: eu = -v - w; uu = eu; break;
                }
                const double r = rr[n];
                if (q == 0) out[n] = r * wq + r * wq * (-1.5 * usq[n]);
                else out[n] = r * wq + r * wq * (3.0 * eu + 4.5 * uu * uu - 1.5 * usq[n]);
            }
        }
    }
    if (p->outlet_bc == 1 && f_convec && g_convec && phi_convec) {
        const long plane = (long)(nx + 2) * (ny + 2);
        for (int j = 0; j <= ny + 1; ++j)
            for (int i = 0; i <= nx + 1; ++i) {
                const long c = i + (long)j * (nx + 2);
                for (int q = 0; q < 19; ++q) {
                    f_convec[c + q * plane] = f[q][F.at(i, j, nz)];
                    g_convec[c + q * plane] = g[q][F.at(i, j, nz)];
                }
                phi_convec[c] = phi[P.at(i, j, nz)];
            }
    }
    return MFLBM_OK;
}

extern "C" int mflbm_inlet_profile(const mflbm_params* p, double uin_avg, double la_x, double la_y, double* w_in) {
    // initialization_open_velocity_inlet_BC (Init_multiphase.F90:194-236) with the rectangular-duct series
    // of inlet_vel_profile_rectangular (Misc.F90:625-665), 1000 terms.  The series divides by n**5 in
    // default (32-bit) integer arithmetic, which wraps for n >= 75; the wrapped divisor is reproduced.
    if (bad_params(p) || !w_in) return MFLBM_ERR_ARG;
    const int nx = p->nx, ny = p->ny;
    const int nterms = 1000;
    const double pi = lat::PI;
    const double pi5 = std::pow(pi, 5.0);          // the reference's compile-time pi**5
    const double pim3 = 1.0 / (pi * pi * pi);      // and pi**(-3)
    for (int j = 0; j <= ny + 1; ++j)
        for (int i = 0; i <= nx + 1; ++i) w_in[i + (long)j * (nx + 2)] = 0.0;
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i)
            if (i > 1 && i < nx && j > 1 && j < ny) w_in[i + (long)j * (nx + 2)] = uin_avg;
    const double a = 0.5 * la_x, b = 0.5 * la_y;
    double tmp1 = 0.0;
    for (int n = 1; n <= nterms; n += 2) {
        uint32_t n5 = (uint32_t)n;
        n5 = n5 * (uint32_t)n * (uint32_t)n * (uint32_t)n * (uint32_t)n;
        tmp1 = tmp1 + std::tanh(0.5 * (double)n * pi * b / a) / (double)(int32_t)n5;
    }
    double tmp2 = 1.0 - 192.0 / pi5 * (a / b) * tmp1;
    tmp2 = -3.0 * uin_avg / (tmp2 * (a * a));
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
            const int x = i, y = j;
            if (x > 1 && x < nx && y > 1 && y < ny) {
                const double xx = x - 1.5 - a;
                const double yy = y - 1.5 - b;
                double tmp3 = 0.0;
                for (int n = 1; n <= nterms; n += 2) {
                    const double n3 = (double)(n * n * n);
                    tmp3 = tmp3 + std::pow(-1.0, 0.5 * (double)(n - 1)) * std::cos(0.5 * n * pi * xx / a) / n3 *
                                      (1.0 - (std::exp(0.5 * n * pi * (yy - b) / a) + std::exp(0.5 * n * pi * (-yy - b) / a)) /
                                                 (1.0 + std::exp(0.5 * n * pi * (-b - b) / a)));
                }
                w_in[i + (long)j * (nx + 2)] = tmp3 * (-16.0 * tmp2 * (a * a) * pim3);
            }
        }
    return MFLBM_OK;
}

extern "C" int mflbm_sp_init_fields(const mflbm_sp_params* p, const double* rho0, const double* u0, const double* v0,
                                    const double* w0, double* const* f, double* f_convec) {
    // initialization_new / initialization_new_pdf (singlephase_3D/0.src/Initialization.F90:207-307): the
    // second-order equilibrium at (rho, u, v, w) over the whole halo-1 array, each population written as
    // the reference writes it (e.g. f7 = rho w2 + rho w2 (3 (u+v) + 4.5 (u+v)(u+v) - 1.5 |u|^2)).
    if (!p || p->nx < 3 || p->ny < 3 || p->nz < 3 || !f) return MFLBM_ERR_ARG;
    for (int q = 0; q < 19; ++q)
        if (!f[q]) return MFLBM_ERR_ARG;
    const Grid G(p->nx, p->ny, p->nz, 1);
    const long n = G.size();
    const double w0e = lat::W0, w1 = lat::W1, w2 = lat::W2;
    for (long t = 0; t < n; ++t) {
        const double u = u0 ? u0[t] : 0.0, v = v0 ? v0[t] : 0.0, w = w0 ? w0[t] : 0.0;
        const double rho1 = rho0 ? rho0[t] : 1.0;
        const double usqrt = u * u + v * v + w * w;
        // axis directions: (component, sign) of e_q . u
        auto axis = [&](double c, double cu) { return rho1 * w1 + rho1 * w1 * (3.0 * cu + 4.5 * c * c - 1.5 * usqrt); };
        auto diag = [&](double cu) { return rho1 * w2 + rho1 * w2 * (3.0 * cu + 4.5 * cu * cu - 1.5 * usqrt); };
        f[0][t] = rho1 * w0e + rho1 * w0e * (-1.5 * usqrt);
        f[1][t] = axis(u, u);
        f[2][t] = axis(u, -u);
        f[3][t] = axis(v, v);
        f[4][t] = axis(v, -v);
        f[5][t] = axis(w, w);
        f[6][t] = axis(w, -w);
        f[7][t] = diag(u + v);
        f[8][t] = diag(-u + v);
        f[9][t] = diag(u - v);
        f[10][t] = diag(-u - v);
        f[11][t] = diag(u + w);
        f[12][t] = diag(-u + w);
        f[13][t] = diag(u - w);
        f[14][t] = diag(-u - w);
        f[15][t] = diag(v + w);
        f[16][t] = diag(-v + w);
        f[17][t] = diag(v - w);
        f[18][t] = diag(-v - w);
    }
    if (f_convec) {
        // f_convec_bc(i, j, q) = f_q(i, j, nz) over the halo-1 plane (Initialization.F90:281-304)
        const long pl = (long)(p->nx + 2) * (p->ny + 2);
        for (int q = 0; q < 19; ++q)
            for (int j = 0; j <= p->ny + 1; ++j)
                for (int i = 0; i <= p->nx + 1; ++i) f_convec[i + (long)j * (p->nx + 2) + q * pl] = f[q][G.at(i, j, p->nz)];
    }
    return MFLBM_OK;
}
