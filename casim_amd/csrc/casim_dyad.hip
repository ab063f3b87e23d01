// casim_dyad.hip -- implementation of include/casim_dyad.h for gfx950.
//
// One handle = one batch of sites resident in HBM.  Device work of a handle
// runs on its own HIP stream; the fused K-step path is captured into a HIP
// graph once per (K, dt, buffer parity) and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/casim_dyad.h"
#include "dyad_internal.hpp"
#include "dyad_kernels.hpp"
#include "mrg32k3a.hpp"

using namespace casim;

namespace casim {
std::string &last_error() {
    thread_local std::string e;
    return e;
}
}  // namespace casim

namespace casim {

// GaussElimination of DyadRDCPU.cpp:163-193 on the equilibrium system of one
// buffer (:53-81): row t: sum_I kon_t C_t B_I + koff_t B_t = kon_t C_t Btot.
std::vector<double> buffer_equilibrium(const std::vector<double> &kon, const std::vector<double> &koff,
                                       const std::vector<double> &c, double total) {
    const int n = (int)kon.size(), m = n;
    std::vector<std::vector<double>> a(n, std::vector<double>(m + 1, 0.0));
    for (int t = 0; t < n; ++t) {
        for (int I = 0; I < m; ++I) a[t][I] = kon[t] * c[t] + (I == t ? koff[t] : 0.0);
        a[t][m] = kon[t] * c[t] * total;
    }
    std::vector<int> where(m, -1);
    for (int col = 0, row = 0; col < m && row < n; ++col) {
        int sel = row;
        for (int i = row; i < n; ++i)
            if (std::fabs(a[i][col]) > std::fabs(a[sel][col])) sel = i;
        if (std::fabs(a[sel][col]) < 1e-10) continue;
        for (int i = col; i <= m; ++i) std::swap(a[sel][i], a[row][i]);
        where[col] = row;
        for (int i = 0; i < n; ++i)
            if (i != row) {
                double q = a[i][col] / a[row][col];
                for (int j = col; j <= m; ++j) a[i][j] -= a[row][j] * q;
            }
        ++row;
    }
    std::vector<double> ans(m, 0.0);
    for (int i = 0; i < m; ++i)
        if (where[i] != -1) ans[i] = a[where[i]][m] / a[where[i]][i];
    return ans;
}

}  // namespace casim

namespace {

void check_seed(const uint64_t *s) {
    for (int i = 0; i < 3; ++i) REQUIRE(s[i] < mrg::M1, CASIM_E_INVALID, "RNG state component >= m1");
    for (int i = 3; i < 6; ++i) REQUIRE(s[i] < mrg::M2, CASIM_E_INVALID, "RNG state component >= m2");
    REQUIRE(s[0] || s[1] || s[2], CASIM_E_INVALID, "RNG state: first three components are 0");
    REQUIRE(s[3] || s[4] || s[5], CASIM_E_INVALID, "RNG state: last three components are 0");
}

template <class T>
T *dalloc(size_t n) {
    void *p = nullptr;
    if (n == 0) n = 1;
    HIPCK(hipMalloc(&p, n * sizeof(T)));
    return (T *)p;
}

}  // namespace


namespace {

template <typename T, int DIM, int NI, int NB>
void step_launch(CasimDyad *d, double dt) {
    const int th = d->threads;
    dim3 grid((d->P.NS + th - 1) / th, d->P.S);
    const T *gin = (const T *)d->g[d->cur];
    T *gout = (T *)d->g[d->cur ^ 1];
    const T *fin = (const T *)d->f[d->cur];
    T *fout = (T *)d->f[d->cur ^ 1];
    hipLaunchKernelGGL((k_step<T, DIM, NI, NB>), grid, dim3(th), 0, d->stream, gin, gout, fin, fout, d->d_off,
                       d->d_snode, d->d_sJ, d->nch, d->P, dt);
}

template <typename T>
void step_dispatch(CasimDyad *d, double dt) {
    const Params &p = d->P;
    if (d->use_sweep && d->sweep_threads > 0) {
        if (sizeof(T) == 8) return p.dim == 3 ? sweep_launch_f64_3d(d, dt) : sweep_launch_f64_2d(d, dt);
        return p.dim == 3 ? sweep_launch_f32_3d(d, dt) : sweep_launch_f32_2d(d, dt);
    }
    if (p.dim == 3) {
        if (p.ni == 1 && p.nb == 3) return step_launch<T, 3, 1, 3>(d, dt);
        if (p.ni == 2 && p.nb == 4) return step_launch<T, 3, 2, 4>(d, dt);
        return step_launch<T, 3, 0, 0>(d, dt);
    }
    if (p.ni == 1 && p.nb == 3) return step_launch<T, 2, 1, 3>(d, dt);
    if (p.ni == 2 && p.nb == 4) return step_launch<T, 2, 2, 4>(d, dt);
    return step_launch<T, 2, 0, 0>(d, dt);
}

}  // namespace

void CasimDyad::launch_step(double dt) {
    if (precision == CASIM_FP64)
        step_dispatch<double>(this, dt);
    else
        step_dispatch<float>(this, dt);
    HIPCK(hipGetLastError());
    cur ^= 1;
}

bool CasimDyad::pair_ok() const {
    // a cytosol pool moves the Dirichlet rim every step: one step per pass
    if (steps_per_pass < 2 || !use_sweep || sweep_threads <= 0 || (Q.on && Q.cyto)) return false;
    // a cytosol box steps between passes only with an even number of substeps
    if (box && casim::box_substeps(this) % 2 != 0) return false;
    return 4 * (size_t)P.n_grid * P.ny * P.nz * tsize <= 64 * 1024;
}

void CasimDyad::launch_pair(double dt) {
    if (precision == CASIM_FP64)
        P.dim == 3 ? sweep2_launch_f64_3d(this, dt) : sweep2_launch_f64_2d(this, dt);
    else
        P.dim == 3 ? sweep2_launch_f32_3d(this, dt) : sweep2_launch_f32_2d(this, dt);
    HIPCK(hipGetLastError());
    cur ^= 1;
}

void CasimDyad::launch_channels(int mode, double dt) {
    const dim3 grid(range_count());
    hipStream_t st = range_st();
    const int s0 = range_s0;
    if (precision == CASIM_FP64) {
        if (chan_small)
            hipLaunchKernelGGL((k_channels<double, CASIM_MODEL_SHEEP_CANNELL, 1, 3, 3>), grid, dim3(64), 0, st,
                               (const double *)g[cur], (const double *)f[cur], Ch, P, Q, mode, dt, s0);
        else
            hipLaunchKernelGGL((k_channels<double, 0, 0, 0, 0>), grid, dim3(64), 0, st, (const double *)g[cur],
                               (const double *)f[cur], Ch, P, Q, mode, dt, s0);
    } else {
        if (chan_small)
            hipLaunchKernelGGL((k_channels<float, CASIM_MODEL_SHEEP_CANNELL, 1, 3, 3>), grid, dim3(64), 0, st,
                               (const float *)g[cur], (const float *)f[cur], Ch, P, Q, mode, dt, s0);
        else
            hipLaunchKernelGGL((k_channels<float, 0, 0, 0, 0>), grid, dim3(64), 0, st, (const float *)g[cur],
                               (const float *)f[cur], Ch, P, Q, mode, dt, s0);
    }
    HIPCK(hipGetLastError());
}

void CasimDyad::launch_rim(const double *ions, const double *bound) {
    dim3 grid((P.NS + 255) / 256, P.S);
    if (precision == CASIM_FP64)
        hipLaunchKernelGGL(k_rim<double>, grid, dim3(256), 0, stream, (double *)g[0], (double *)g[1],
                           (double *)f[0], (double *)f[1], ions, bound, P);
    else
        hipLaunchKernelGGL(k_rim<float>, grid, dim3(256), 0, stream, (float *)g[0], (float *)g[1], (float *)f[0],
                           (float *)f[1], ions, bound, P);
    HIPCK(hipGetLastError());
}

void CasimDyad::launch_efflux() {
    dim3 grid(P.S * (P.ni + P.nb));
    if (precision == CASIM_FP64)
        hipLaunchKernelGGL((k_efflux<double>), grid, dim3(64), 0, stream, (const double *)g[cur], d_eff, P,
                           grad_mult, sxm, sym);
    else
        hipLaunchKernelGGL((k_efflux<float>), grid, dim3(64), 0, stream, (const float *)g[cur], d_eff, P,
                           grad_mult, sxm, sym);
    HIPCK(hipGetLastError());
}

extern "C" {

const char *casim_last_error(void) { return casim::last_error().c_str(); }

const char *casim_version(void) { return "casim_dyad 0.1 gfx950"; }

int casim_rng_stream_states(const uint64_t seed[6], int32_t first, int32_t n, uint64_t *out) {
    return guard([&] {
        REQUIRE(seed && out && first >= 0 && n >= 0, CASIM_E_INVALID, "bad arguments");
        check_seed(seed);
        uint32_t A1[9], A2[9], J1[9], J2[9];
        mrg::base_matrices(A1, A2);
        mrg::two_pow(A1, 127, J1, mrg::C1, mrg::M1);
        mrg::two_pow(A2, 127, J2, mrg::C2, mrg::M2);
        uint32_t s[6];
        for (int i = 0; i < 6; ++i) s[i] = (uint32_t)seed[i];
        for (int k = 0; k < first + n; ++k) {
            if (k >= first)
                for (int i = 0; i < 6; ++i) out[6 * (size_t)(k - first) + i] = s[i];
            mrg::matvec(J1, s, s, mrg::C1, mrg::M1);
            mrg::matvec(J2, s + 3, s + 3, mrg::C2, mrg::M2);
        }
    });
}

int casim_rng_advance(uint64_t state[6], int64_t n_draws) {
    return guard([&] {
        REQUIRE(state && n_draws >= 0, CASIM_E_INVALID, "bad arguments");
        check_seed(state);
        uint32_t A1[9], A2[9], P1[9], P2[9], s[6];
        mrg::base_matrices(A1, A2);
        mrg::power(A1, 2 * (uint64_t)n_draws, P1, mrg::C1, mrg::M1);
        mrg::power(A2, 2 * (uint64_t)n_draws, P2, mrg::C2, mrg::M2);
        for (int i = 0; i < 6; ++i) s[i] = (uint32_t)state[i];
        mrg::matvec(P1, s, s, mrg::C1, mrg::M1);
        mrg::matvec(P2, s + 3, s + 3, mrg::C2, mrg::M2);
        for (int i = 0; i < 6; ++i) state[i] = s[i];
    });
}

int casim_rng_draws(uint64_t state[6], int32_t n_draws, double *out) {
    return guard([&] {
        REQUIRE(state && out && n_draws >= 0, CASIM_E_INVALID, "bad arguments");
        check_seed(state);
        uint32_t s[6];
        for (int i = 0; i < 6; ++i) s[i] = (uint32_t)state[i];
        for (int i = 0; i < n_draws; ++i) out[i] = mrg::draw(s);
        for (int i = 0; i < 6; ++i) state[i] = s[i];
    });
}

int casim_dyad_create(const CasimDyadConfig *c, CasimDyad **out) {
    CasimDyad *d = nullptr;
    int rc = guard([&] {
        REQUIRE(c && out, CASIM_E_INVALID, "null argument");
        *out = nullptr;
        REQUIRE(c->dim == 2 || c->dim == 3, CASIM_E_INVALID, "dim must be 2 or 3");
        REQUIRE(c->nx >= 5 && c->ny >= 5, CASIM_E_INVALID, "grid needs nx, ny >= 5");
        // the sweeps address a site's field with a 32-bit byte offset
        REQUIRE((double)c->nx * c->ny * (c->dim == 3 ? c->nz : 1) * 8.0 < 2147483648.0, CASIM_E_INVALID,
                "a site's grid must stay below 2^28 nodes");
        REQUIRE(c->dim == 2 || c->nz >= 2, CASIM_E_INVALID, "3D grid needs nz >= 2");
        REQUIRE(c->n_sites >= 1 && c->n_sites <= 65535, CASIM_E_INVALID, "n_sites must be in 1..65535");
        REQUIRE(c->n_ions >= 1 && c->n_ions <= MAXI, CASIM_E_INVALID, "n_ions out of range (1..4)");
        REQUIRE(c->n_buffers >= 0 && c->n_buffers <= MAXU, CASIM_E_INVALID, "n_buffers out of range (0..8)");
        REQUIRE(c->n_bound >= 0 && c->n_bound <= MAXB, CASIM_E_INVALID, "n_bound out of range (0..8)");
        REQUIRE(c->n_channel_types >= 1 && c->n_channel_types <= MAXT, CASIM_E_INVALID,
                "n_channel_types out of range (1..4)");
        REQUIRE(c->precision == CASIM_FP64 || c->precision == CASIM_FP32, CASIM_E_INVALID, "bad precision");
        REQUIRE(c->dx > 0 && c->dy > 0 && (c->dim == 2 || c->dz > 0), CASIM_E_INVALID, "spacing must be > 0");
        d = new CasimDyad();
        HIPCK(hipGetDevice(&d->dev));
        Params &P = d->P;
        P.dim = c->dim;
        P.nx = c->nx;
        P.ny = c->ny;
        P.nz = c->dim == 3 ? c->nz : 1;
        P.S = c->n_sites;
        P.NS = P.nx * P.ny * P.nz;
        P.nFP = P.nx * P.ny;
        P.ni = c->n_ions;
        P.nb = c->n_bound;
        P.nu = c->n_buffers;
        for (int j = 0; j < P.ni; ++j) {
            P.Dion[j] = c->ion_D[j];
            P.c0[j] = c->ion_c0[j];
            P.sl[j] = c->sl_binds ? c->sl_binds[j] : 0;
            if (P.sl[j]) {
                P.N1[j] = c->sl_N1[j];
                P.K1[j] = c->sl_K1[j];
                P.N2[j] = c->sl_N2[j];
                P.K2[j] = c->sl_K2[j];
            }
        }
        for (int u = 0; u < P.nu; ++u) {
            P.total[u] = c->buffer_total[u];
            P.ufaces[u] = (c->buffer_faces_only && c->dim == 3) ? c->buffer_faces_only[u] : 0;
        }
        // bound fields must be ion-major over (ion, buffer) with no repeats
        int prev = -1;
        for (int b = 0; b < P.nb; ++b) {
            int key = c->bound_ion[b] * MAXU + c->bound_buffer[b];
            REQUIRE(c->bound_ion[b] >= 0 && c->bound_ion[b] < P.ni && c->bound_buffer[b] >= 0 &&
                        c->bound_buffer[b] < P.nu,
                    CASIM_E_INVALID, "bound field refers to a missing ion or buffer");
            REQUIRE(key > prev, CASIM_E_INVALID, "bound fields must be ion-major (DyadRDCPU order)");
            prev = key;
            P.bion[b] = c->bound_ion[b];
            P.bbuf[b] = c->bound_buffer[b];
            P.Db[b] = c->bound_D[b];
            P.kon[b] = c->bound_kon[b];
            P.koff[b] = c->bound_koff[b];
            P.bface[b] = P.ufaces[P.bbuf[b]];
            REQUIRE(!P.bface[b] || P.Db[b] == 0.0, CASIM_E_INVALID, "membrane-only buffers must have D = 0");
        }
        P.n_grid = P.ni;
        P.n_face = 0;
        for (int b = 0; b < P.nb; ++b) P.bslot[b] = P.bface[b] ? P.n_face++ : P.n_grid++;
        for (int j = 0; j < P.ni; ++j) P.gD[j] = P.Dion[j];
        for (int b = 0; b < P.nb; ++b)
            if (!P.bface[b]) P.gD[P.bslot[b]] = P.Db[b];
        // equilibrium (DyadRDCPU.cpp:53-81)
        for (int b = 0; b < P.nb; ++b) P.binit[b] = 0.0;
        if (c->start_equilibrium) {
            for (int u = 0; u < P.nu; ++u) {
                std::vector<double> kon, koff, cc;
                std::vector<int> idx;
                for (int b = 0; b < P.nb; ++b)
                    if (P.bbuf[b] == u) {
                        kon.push_back(P.kon[b]);
                        koff.push_back(P.koff[b]);
                        cc.push_back(P.c0[P.bion[b]]);
                        idx.push_back(b);
                    }
                if (idx.empty()) continue;
                auto sol = buffer_equilibrium(kon, koff, cc, P.total[u]);
                for (size_t t = 0; t < idx.size(); ++t) P.binit[idx[t]] = sol[t];
            }
        }
        d->dx = c->dx;
        d->dy = c->dy;
        d->dz = c->dim == 3 ? c->dz : 1e-5;
        d->zext = c->z;
        d->radius = c->radius;
        d->voltage = c->voltage;
        P.ihx2 = 1.0 / (c->dx * c->dx);
        P.ihy2 = 1.0 / (c->dy * c->dy);
        P.ihz2 = c->dim == 3 ? 1.0 / (c->dz * c->dz) : 0.0;
        P.cc = 2.0 * (P.ihx2 + P.ihy2 + P.ihz2);
        d->elem_volume = c->dim == 3 ? 0.5 * c->dx * c->dy * c->dz : c->dx * c->dy * c->z;
        d->grad_mult = 3.0 / (4 * std::acos(-1.0) * c->radius * c->radius * c->radius);
        d->sxm = c->dim == 3 ? c->dy * c->dz / c->dx : c->dy * c->z / c->dx;
        d->sym = c->dim == 3 ? c->dx * c->dz / c->dy : c->dx * c->z / c->dy;
        d->precision = c->precision;
        d->tsize = c->precision == CASIM_FP64 ? 8 : 4;
        d->start_equilibrium = c->start_equilibrium != 0;
        {
            // plane-marching sweep: one plane of every grid field in LDS (x2)
            const int Pl = P.ny * P.nz;
            const size_t lds = 2 * (size_t)P.n_grid * Pl * d->tsize;
            const int th = Pl > 256 ? 256 : (Pl + 63) / 64 * 64;
            if (Pl <= 4 * 512 && lds <= 65536 && (Pl + th - 1) / th <= 4) {
                d->sweep_threads = th;
                d->sweep_chunk = P.nx - 2;
            }
        }

        // channels
        Channels &Ch = d->Ch;
        d->h_off.assign(c->site_channel_offset, c->site_channel_offset + P.S + 1);
        REQUIRE(d->h_off[0] == 0, CASIM_E_INVALID, "site_channel_offset[0] must be 0");
        d->M = 0;
        for (int s = 0; s < P.S; ++s) {
            int n = d->h_off[s + 1] - d->h_off[s];
            REQUIRE(n >= 0 && n <= MAX_SITE_CHANNELS, CASIM_E_INVALID, "a site has more than 127 channels");
            d->M = std::max(d->M, n);
        }
        d->nch = d->h_off[P.S];
        d->h_node.resize(d->nch);
        std::vector<int> types(d->nch);
        for (int s = 0; s < P.S; ++s)
            for (int k = d->h_off[s]; k < d->h_off[s + 1]; ++k) {
                // Point::GetCoordsOnGrid (Point.cpp:17-23), z = 0 (2D maps with dz = 1e-5)
                int gx = (int)((c->channel_x[k] + 0.5 * c->dx) / c->dx);
                int gy = (int)((c->channel_y[k] + 0.5 * c->dy) / c->dy);
                int gz = (int)((0.0 + 0.5 * d->dz) / d->dz);
                REQUIRE(gx >= 1 && gx <= P.nx - 2 && gy >= 1 && gy <= P.ny - 2 && gz < P.nz, CASIM_E_INVALID,
                        "a channel mouth lies on or outside the Dirichlet rim");
                d->h_node[k] = gz + P.nz * (gy + P.ny * gx);
                types[k] = c->channel_type[k];
                REQUIRE(types[k] >= 0 && types[k] < c->n_channel_types, CASIM_E_INVALID, "bad channel type");
            }
        // node-sorted list per site (stable: the last channel on a node wins)
        d->h_perm.resize(d->nch);
        d->h_snode.resize(d->nch);
        d->h_sperm.resize(d->nch);
        for (int s = 0; s < P.S; ++s) {
            int o = d->h_off[s], n = d->h_off[s + 1] - o;
            std::vector<int> ord(n);
            std::iota(ord.begin(), ord.end(), 0);
            std::stable_sort(ord.begin(), ord.end(),
                             [&](int a, int b) { return d->h_node[o + a] < d->h_node[o + b]; });
            for (int r = 0; r < n; ++r) {
                d->h_perm[o + ord[r]] = r;
                d->h_snode[o + r] = d->h_node[o + ord[r]];
                d->h_sperm[o + r] = ord[r];
            }
        }
        for (int t = 0; t < c->n_channel_types; ++t) {
            int m = c->type_model[t];
            REQUIRE(m >= 1 && m <= 3, CASIM_E_INVALID, "unknown kinetic model");
            Ch.model[t] = m;
            for (int q = 0; q < NPAR; ++q) Ch.par[t * NPAR + q] = c->type_params[t * CASIM_MAX_MODEL_PARAMS + q];
            for (int k = 0; k < MAXI; ++k) Ch.g[t * MAXI + k] = k < P.ni ? c->type_conductance[t * P.ni + k] : 0.0;
        }
        // the channel kernel's compile-time form: every type SheepCannell with one
        // ion, three bound fields and three buffers, or the general one
        bool sheep = true;
        for (int t = 0; t < c->n_channel_types; ++t) sheep = sheep && Ch.model[t] == CASIM_MODEL_SHEEP_CANNELL;
        d->chan_small = sheep && P.ni == 1 && P.nb == 3 && P.nu == 3;
        Ch.M = d->M;
        Ch.nch = d->nch;
        REQUIRE(c->calcium_index >= -1 && c->calcium_index < P.ni && c->magnesium_index >= -1 &&
                    c->magnesium_index < P.ni,
                CASIM_E_INVALID, "calcium/magnesium ion index out of range");
        Ch.ca = c->calcium_index;
        Ch.mg = c->magnesium_index;
        d->h_rng0.resize((size_t)P.S * 6);
        for (int s = 0; s < P.S; ++s) {
            check_seed(c->rng_state + 6 * (size_t)s);
            for (int i = 0; i < 6; ++i) d->h_rng0[6 * (size_t)s + i] = (uint32_t)c->rng_state[6 * (size_t)s + i];
        }
        // jump table A^(2p), p < JUMP_N
        std::vector<uint32_t> jump((size_t)JUMP_N * 18);
        {
            uint32_t A1[9], A2[9], S1[9], S2[9], P1[9], P2[9];
            mrg::base_matrices(A1, A2);
            mrg::matmul(A1, A1, S1, mrg::C1, mrg::M1);
            mrg::matmul(A2, A2, S2, mrg::C2, mrg::M2);
            mrg::identity(P1);
            mrg::identity(P2);
            for (int p = 0; p < JUMP_N; ++p) {
                std::memcpy(&jump[(size_t)p * 18], P1, sizeof P1);
                std::memcpy(&jump[(size_t)p * 18 + 9], P2, sizeof P2);
                mrg::matmul(S1, P1, P1, mrg::C1, mrg::M1);
                mrg::matmul(S2, P2, P2, mrg::C2, mrg::M2);
            }
        }

        HIPCK(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
        HIPCK(hipStreamCreateWithFlags(&d->stream2, hipStreamNonBlocking));
        HIPCK(hipEventCreateWithFlags(&d->ev_fork, hipEventDisableTiming));
        HIPCK(hipEventCreateWithFlags(&d->ev_off, hipEventDisableTiming));
        HIPCK(hipEventCreateWithFlags(&d->ev_join, hipEventDisableTiming));
        for (int k = 0; k < 2; ++k) {
            HIPCK(hipMalloc(&d->g[k], std::max<size_t>(1, d->grid_elems()) * d->tsize));
            HIPCK(hipMalloc(&d->f[k], std::max<size_t>(1, d->face_elems()) * d->tsize));
        }
        const int K = P.ni + P.nb + P.nu;
        d->d_off = dalloc<int>(P.S + 1);
        d->d_node = dalloc<int>(d->nch);
        d->d_type = dalloc<int>(d->nch);
        d->d_perm = dalloc<int>(d->nch);
        d->d_snode = dalloc<int>(d->nch);
        d->d_sperm = dalloc<int>(d->nch);
        d->d_state = dalloc<int>(d->nch);
        d->d_chans = dalloc<int>(P.S);
        d->d_rng = dalloc<uint32_t>((size_t)P.S * 6);
        d->d_rng0 = dalloc<uint32_t>((size_t)P.S * 6);
        d->d_jump = dalloc<uint32_t>(jump.size());
        d->d_mouths = dalloc<double>((size_t)K * P.S * std::max(1, d->M));
        d->d_jsr = dalloc<double>((size_t)P.ni * P.S * std::max(1, d->M));
        d->d_sJ = dalloc<double>((size_t)P.ni * std::max(1, d->nch));
        d->d_snap = dalloc<double>((size_t)(P.n_grid + P.n_face) * std::max(1, d->nch));
        d->d_pg = dalloc<double>((size_t)P.ni * std::max(1, d->nch));
        d->d_pjv = dalloc<double>((size_t)P.ni * std::max(1, d->nch));
        d->d_totals = dalloc<double>((size_t)P.ni * P.S);
        d->d_rim_ions = dalloc<double>((size_t)P.ni * P.S);
        d->d_rim_bound = dalloc<double>((size_t)std::max(1, P.nb) * P.S);
        d->d_eff = dalloc<double>((size_t)(P.ni + P.nb) * P.S);
        d->d_stats = dalloc<double>((size_t)NSTAT * P.S);
        auto up = [&](void *dst, const void *src, size_t bytes) {
            if (bytes) HIPCK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, d->stream));
        };
        up(d->d_off, d->h_off.data(), sizeof(int) * (P.S + 1));
        up(d->d_node, d->h_node.data(), sizeof(int) * d->nch);
        up(d->d_type, types.data(), sizeof(int) * d->nch);
        up(d->d_perm, d->h_perm.data(), sizeof(int) * d->nch);
        up(d->d_snode, d->h_snode.data(), sizeof(int) * d->nch);
        up(d->d_sperm, d->h_sperm.data(), sizeof(int) * d->nch);
        up(d->d_rng0, d->h_rng0.data(), sizeof(uint32_t) * P.S * 6);
        up(d->d_jump, jump.data(), sizeof(uint32_t) * jump.size());
        HIPCK(hipMemsetAsync(d->d_sJ, 0, sizeof(double) * P.ni * std::max(1, d->nch), d->stream));
        HIPCK(hipMemsetAsync(d->d_totals, 0, sizeof(double) * P.ni * P.S, d->stream));
        HIPCK(hipMemsetAsync(d->d_mouths, 0, sizeof(double) * K * P.S * std::max(1, d->M), d->stream));
        HIPCK(hipMemsetAsync(d->d_eff, 0, sizeof(double) * (P.ni + P.nb) * P.S, d->stream));
        Ch.off = d->d_off;
        Ch.node = d->d_node;
        Ch.type = d->d_type;
        Ch.perm = d->d_perm;
        Ch.snode = d->d_snode;
        Ch.sperm = d->d_sperm;
        Ch.state = d->d_state;
        Ch.rng = d->d_rng;
        Ch.jump = d->d_jump;
        Ch.mouths = d->d_mouths;
        Ch.jsr = d->d_jsr;
        Ch.sJ = d->d_sJ;
        Ch.snap = d->d_snap;
        Ch.pg = d->d_pg;
        Ch.pjv = d->d_pjv;
        Ch.totals = d->d_totals;
        Ch.stats = d->d_stats;
        Ch.trace = nullptr;
        Ch.trace_every = 1;
        Ch.trace_bins = 0;
        HIPCK(hipStreamSynchronize(d->stream));
        if (casim_dyad_reset(d) != CASIM_OK) throw Err(CASIM_E_DEVICE, casim::last_error());
        *out = d;
    });
    if (rc != CASIM_OK && d) {
        std::string keep = casim::last_error();
        casim_dyad_destroy(d);
        casim::last_error() = keep;
    }
    return rc;
}

int casim_dyad_destroy(CasimDyad *d) {
    if (!d) return CASIM_OK;
    (void)hipSetDevice(d->dev);
    if (d->stream) (void)hipStreamSynchronize(d->stream);
    if (d->stream2) (void)hipStreamSynchronize(d->stream2);
    d->clear_graphs();
    casim::box_free(d);
    void *ptrs[] = {d->g[0], d->g[1], d->f[0], d->f[1], d->d_off, d->d_node, d->d_type, d->d_perm, d->d_snode,
                    d->d_sperm, d->d_state, d->d_chans, d->d_rng, d->d_rng0, d->d_jump, d->d_mouths, d->d_jsr, d->d_sJ,
                    d->d_snap, d->d_pg, d->d_pjv, d->d_totals, d->d_rim_ions, d->d_rim_bound, d->d_eff, d->d_stats, d->Ch.trace};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (double *p : d->pool_buffers)
        if (p) (void)hipFree(p);
    if (d->stream) (void)hipStreamDestroy(d->stream);
    if (d->stream2) (void)hipStreamDestroy(d->stream2);
    for (hipEvent_t e : {d->ev_fork, d->ev_off, d->ev_join})
        if (e) (void)hipEventDestroy(e);
    delete d;
    return CASIM_OK;
}

int casim_dyad_reset(CasimDyad *d) {
    return guard([&] {
        REQUIRE(d, CASIM_E_INVALID, "null handle");
        d->use();
        const Params &P = d->P;
        size_t n = std::max((size_t)P.S * P.NS, (size_t)P.S * 2 * P.nFP);
        dim3 grid((unsigned)((n + 255) / 256));
        if (d->precision == CASIM_FP64)
            hipLaunchKernelGGL(k_reset<double>, grid, dim3(256), 0, d->stream, (double *)d->g[0], (double *)d->g[1],
                               (double *)d->f[0], (double *)d->f[1], P);
        else
            hipLaunchKernelGGL(k_reset<float>, grid, dim3(256), 0, d->stream, (float *)d->g[0], (float *)d->g[1],
                               (float *)d->f[0], (float *)d->f[1], P);
        HIPCK(hipGetLastError());
        HIPCK(hipMemsetAsync(d->d_state, 0, sizeof(int) * std::max(1, d->nch), d->stream));
        HIPCK(hipMemsetAsync(d->d_stats, 0, sizeof(double) * NSTAT * P.S, d->stream));
        if (d->Ch.trace)
            HIPCK(hipMemsetAsync(d->Ch.trace, 0xFF, (size_t)d->Ch.trace_bins * P.S, d->stream));
        if (d->Q.on) {
            hipLaunchKernelGGL(k_pools_reset, dim3((P.S + 127) / 128), dim3(128), 0, d->stream, d->Q, P.S, P.ni,
                               d->d_jc0);
            HIPCK(hipGetLastError());
        }
        HIPCK(hipMemcpyAsync(d->d_rng, d->d_rng0, sizeof(uint32_t) * P.S * 6, hipMemcpyDeviceToDevice, d->stream));
        HIPCK(hipStreamSynchronize(d->stream));
        if (d->box) casim::box_reset(d);
        d->cur = 0;
        d->have_update = false;
    });
}

int casim_dyad_init_opening(CasimDyad *d, int32_t site, int32_t channel, const int32_t *channels) {
    return guard([&] {
        REQUIRE(d, CASIM_E_INVALID, "null handle");
        REQUIRE(site >= -1 && site < d->P.S, CASIM_E_RANGE, "site out of range");
        d->use();
        const int *dch = nullptr;
        if (channels) {
            REQUIRE(site == -1, CASIM_E_INVALID, "channels[] needs site = -1");
            for (int s = 0; s < d->P.S; ++s)
                REQUIRE(channels[s] >= 0 && channels[s] < d->h_off[s + 1] - d->h_off[s], CASIM_E_RANGE,
                        "channel index out of range");
            HIPCK(hipMemcpyAsync(d->d_chans, channels, sizeof(int) * d->P.S, hipMemcpyHostToDevice, d->stream));
            dch = d->d_chans;
        } else {
            for (int s = 0; s < d->P.S; ++s)
                if (site == -1 || s == site)
                    REQUIRE(channel >= 0 && channel < d->h_off[s + 1] - d->h_off[s], CASIM_E_RANGE,
                            "channel index out of range");
        }
        hipLaunchKernelGGL(k_init_opening, dim3((d->P.S + 127) / 128), dim3(128), 0, d->stream, d->Ch, d->P.S, site,
                           channel, dch);
        HIPCK(hipGetLastError());
        HIPCK(hipStreamSynchronize(d->stream));
    });
}

int casim_dyad_update(CasimDyad *d, const double *rim_ions, const double *rim_bound, const double *jsr,
                      double voltage) {
    return guard([&] {
        REQUIRE(d, CASIM_E_INVALID, "null handle");
        const bool pool_rim = d->Q.on && d->Q.cyto;
        const bool box_rim = d->box != nullptr;
        REQUIRE(pool_rim || box_rim || (rim_ions && (rim_bound || d->P.nb == 0)), CASIM_E_INVALID,
                "null rim argument");
        REQUIRE(d->Q.on || jsr, CASIM_E_INVALID, "null jsr argument");
        d->use();
        const Params &P = d->P;
        d->voltage = voltage;
        if (!pool_rim && !box_rim) {
            HIPCK(hipMemcpyAsync(d->d_rim_ions, rim_ions, sizeof(double) * P.ni * P.S, hipMemcpyHostToDevice,
                                 d->stream));
            if (P.nb)
                HIPCK(hipMemcpyAsync(d->d_rim_bound, rim_bound, sizeof(double) * P.nb * P.S, hipMemcpyHostToDevice,
                                     d->stream));
        }
        if (!d->Q.on && d->M)
            HIPCK(hipMemcpyAsync(d->d_jsr, jsr, sizeof(double) * P.ni * P.S * d->M, hipMemcpyHostToDevice, d->stream));
        if (pool_rim)
            d->launch_rim(d->Q.cions, d->Q.cbb);
        else if (box_rim)
            casim::box_rim(d);
        else
            d->launch_rim(d->d_rim_ions, d->d_rim_bound);
        d->launch_efflux();
        d->launch_channels(CH_SNAPSHOT | CH_CURRENTS | CH_ACCUMULATE | CH_POOL_UPDATE, 0.0);
        d->clear_graphs();
        HIPCK(hipStreamSynchronize(d->stream));
        d->have_update = true;
    });
}

}  // extern "C"

namespace {
void copy_out(CasimDyad *d, double *out, const double *src, size_t n) {
    HIPCK(hipMemcpyAsync(out, src, sizeof(double) * n, hipMemcpyDeviceToHost, d->stream));
    HIPCK(hipStreamSynchronize(d->stream));
}
}  // namespace

extern "C" {

int casim_dyad_attach_pools(CasimDyad *d, const CasimPoolConfig *pc) {
    return guard([&] {
        REQUIRE(d && pc, CASIM_E_INVALID, "null argument");
        REQUIRE(!d->Q.on, CASIM_E_STATE, "pools are already attached");
        REQUIRE(!(d->box && pc->cytosol_pool), CASIM_E_STATE, "a cytosol box is attached");
        d->use();
        const Params &P = d->P;
        Pools Q{};
        const int ni = P.ni, S = P.S;
        Q.jsr_const = pc->jsr_constant;
        Q.cq = pc->jsr_cq_total;
        Q.Vjsr = pc->jsr_volume;
        REQUIRE(pc->jsr_constant || pc->jsr_volume > 0, CASIM_E_INVALID, "jSR volume must be > 0");
        Q.Velem = d->elem_volume;
        std::vector<double> jc0(ni), nsr((size_t)ni * S), jnsr((size_t)ni * S);
        for (int i = 0; i < ni; ++i) {
            Q.nCQ[i] = pc->jsr_cq_sites[i];
            Q.KCQ[i] = pc->jsr_k_cq[i];
            Q.tau[i] = pc->jsr_tau_refill[i];
            REQUIRE(pc->jsr_constant || Q.tau[i] != 0.0, CASIM_E_INVALID, "jSR refill time must be non-zero");
            jc0[i] = pc->jsr_c0[i];
            for (int s = 0; s < S; ++s) {
                nsr[s + (size_t)i * S] = pc->nsr_c[i];
                jnsr[s + (size_t)i * S] = jc0[i];  // SRPool's ctor (SRPool.cpp:67-72)
            }
        }
        Q.cyto = pc->cytosol_pool;
        const int nc = pc->cytosol_pool ? pc->n_cyto_buffers : 0;
        REQUIRE(nc >= 0 && P.nu + nc <= MAXCU, CASIM_E_INVALID, "too many cytosol buffers");
        Q.nbc = P.nu + nc;
        int k = 0;
        double Dk[MAXCB];
        for (int j = 0; j < ni; ++j) {
            for (int u = 0; u < Q.nbc; ++u) {
                Q.cbind[j * Q.nbc + u] = -1;
                double kon, koff, D;
                int dy;
                if (u < P.nu) {
                    int b = -1;
                    for (int bb = 0; bb < P.nb; ++bb)
                        if (P.bion[bb] == j && P.bbuf[bb] == u) b = bb;
                    if (b < 0) continue;
                    kon = P.kon[b];
                    koff = P.koff[b];
                    D = P.Db[b];
                    dy = 1;
                } else {
                    const int c = u - P.nu;
                    if (!pc->cyto_buffer_binds[c * ni + j]) continue;
                    kon = pc->cyto_buffer_kon[c * ni + j];
                    koff = pc->cyto_buffer_koff[c * ni + j];
                    D = pc->cyto_buffer_D[c * ni + j];
                    dy = 0;
                }
                REQUIRE(k < MAXCB, CASIM_E_INVALID, "too many cytosol bound fields");
                Q.cbind[j * Q.nbc + u] = k;
                Q.ckon[k] = kon;
                Q.ckoff[k] = koff;
                Q.cb_dyad[k] = dy;
                Dk[k] = D;
                Q.conet_b[k] = pc->cytosol_pool ? pc->cyto_leak_mult * D : 0.0;
                ++k;
            }
            Q.cbg[j] = P.c0[j];
            Q.conet_i[j] = pc->cytosol_pool ? pc->cyto_leak_mult * P.Dion[j] : 0.0;
            Q.serca[j] = (pc->cytosol_pool && pc->serca_binds) ? pc->serca_binds[j] : 0;
            Q.jmax[j] = Q.serca[j] ? pc->serca_jmax[j] : 0.0;
            Q.kup[j] = Q.serca[j] ? pc->serca_kup[j] : 0.0;
        }
        (void)Dk;
        Q.nBc = k;
        for (int u = 0; u < Q.nbc; ++u) Q.ctot[u] = u < P.nu ? P.total[u] : pc->cyto_buffer_total[u - P.nu];
        if (d->start_equilibrium) {
            for (int u = 0; u < Q.nbc; ++u) {
                std::vector<double> kon, koff, cc;
                std::vector<int> idx;
                for (int j = 0; j < ni; ++j) {
                    const int b = Q.cbind[j * Q.nbc + u];
                    if (b < 0) continue;
                    kon.push_back(Q.ckon[b]);
                    koff.push_back(Q.ckoff[b]);
                    cc.push_back(Q.cbg[j]);
                    idx.push_back(b);
                }
                if (idx.empty()) continue;
                auto sol = buffer_equilibrium(kon, koff, cc, Q.ctot[u]);
                for (size_t t = 0; t < idx.size(); ++t) Q.cinit[idx[t]] = sol[t];
            }
        }
        auto buf = [&](size_t n) {
            double *p = dalloc<double>(n);
            d->pool_buffers.push_back(p);
            HIPCK(hipMemsetAsync(p, 0, sizeof(double) * std::max<size_t>(n, 1), d->stream));
            return p;
        };
        Q.jions = buf((size_t)ni * S);
        Q.jnew = buf((size_t)ni * S);
        Q.IjSR = buf((size_t)ni * S);
        Q.jnsr = buf((size_t)ni * S);
        double *dn = buf((size_t)ni * S);
        Q.nsr = dn;
        Q.cions = buf((size_t)ni * S);
        Q.cbuf = buf((size_t)std::max(1, Q.nBc) * S);
        Q.cfree = buf((size_t)std::max(1, Q.nbc) * S);
        Q.cbb = buf((size_t)std::max(1, P.nb) * S);
        d->d_jc0 = buf(ni);
        HIPCK(hipMemcpyAsync(dn, nsr.data(), sizeof(double) * nsr.size(), hipMemcpyHostToDevice, d->stream));
        HIPCK(hipMemcpyAsync(Q.jnsr, jnsr.data(), sizeof(double) * jnsr.size(), hipMemcpyHostToDevice, d->stream));
        HIPCK(hipMemcpyAsync(d->d_jc0, jc0.data(), sizeof(double) * ni, hipMemcpyHostToDevice, d->stream));
        Q.eff = d->d_eff;
        Q.on = 1;
        d->Q = Q;
        hipLaunchKernelGGL(k_pools_reset, dim3((S + 127) / 128), dim3(128), 0, d->stream, d->Q, S, ni, d->d_jc0);
        HIPCK(hipGetLastError());
        HIPCK(hipStreamSynchronize(d->stream));
        d->have_update = false;
        d->clear_graphs();
    });
}

int casim_dyad_get_pools(CasimDyad *d, double *jsr, double *cyto_ions, double *cyto_bound) {
    return guard([&] {
        REQUIRE(d, CASIM_E_INVALID, "null handle");
        REQUIRE(d->Q.on, CASIM_E_STATE, "no pools attached");
        d->use();
        const size_t n = (size_t)d->P.ni * d->P.S;
        if (jsr) copy_out(d, jsr, d->Q.jnew, n);
        if (cyto_ions) copy_out(d, cyto_ions, d->Q.cions, n);
        if (cyto_bound && d->Q.nBc) copy_out(d, cyto_bound, d->Q.cbuf, (size_t)d->Q.nBc * d->P.S);
    });
}

int32_t casim_dyad_cytosol_bound(const CasimDyad *d) { return d && d->Q.on ? d->Q.nBc : 0; }

int casim_dyad_run_rd(CasimDyad *d, double dt, int32_t site) {
    return guard([&] {
        REQUIRE(d, CASIM_E_INVALID, "null handle");
        REQUIRE(site == -1, CASIM_E_INVALID, "the device path advances every site at once (site = -1)");
        REQUIRE(d->have_update, CASIM_E_STATE, "run_rd before the first update");
        REQUIRE(!d->box, CASIM_E_STATE, "with a cytosol box, advance with casim_dyad_step");
        d->use();
        d->launch_step(dt);
        HIPCK(hipStreamSynchronize(d->stream));
    });
}

int casim_dyad_run_mc(CasimDyad *d, double dt, int32_t site) {
    return guard([&] {
        REQUIRE(d, CASIM_E_INVALID, "null handle");
        REQUIRE(site == -1, CASIM_E_INVALID, "the device path gates every site at once (site = -1)");
        REQUIRE(d->have_update, CASIM_E_STATE, "run_mc before the first update");
        d->use();
        d->launch_channels(CH_GATE_STORED, dt);
        HIPCK(hipStreamSynchronize(d->stream));
    });
}

int casim_dyad_step(CasimDyad *d, double dt, int32_t k_steps) {
    return guard([&] {
        REQUIRE(d, CASIM_E_INVALID, "null handle");
        REQUIRE(k_steps >= 1, CASIM_E_INVALID, "k_steps must be >= 1");
        REQUIRE(d->have_update, CASIM_E_STATE, "step before the first update");
        const int M = casim::box_substeps(d);
        REQUIRE(!d->box || k_steps % M == 0, CASIM_E_INVALID,
                "with a cytosol box, k_steps must be a multiple of its substeps");
        d->use();
        // One captured graph covers at most 2048 steps, a whole number of
        // box steps: a longer call replays several graphs, which gives the
        // results of one (each graph ends with the update that the next
        // graph's first channel call expects) while the node count of any
        // one graph, about three per two steps and half-batch, stays bounded.
        const int unit = 2 * M, cap = std::max(unit, 2048 / unit * unit);
        for (int32_t left = k_steps; left > 0;) {
            const int32_t kc = std::min(left, (int32_t)cap);
            left -= kc;
            const int cur0 = d->cur, bcur0 = casim::box_cur(d);
            auto key = std::make_tuple((int)kc, dt, cur0 | bcur0 << 1);
            auto it = d->graphs.find(key);
            if (it == d->graphs.end()) {
                HIPCK(hipStreamBeginCapture(d->stream, hipStreamCaptureModeThreadLocal));
                const bool pool_rim = d->Q.on && d->Q.cyto;
                const bool pairs = d->pair_ok();
                try {
                    int k = 0;
                    if (d->split && pairs && !pool_rim && d->P.S >= 2 && kc >= 2) {
                        // Split pipelining: the two halves of the batch step on two
                        // streams; the second starts one channel call after the
                        // first, so one half's gating runs beside the other half's
                        // pass.  Sites are independent (a cytosol box too: each
                        // site's box takes only that site's effluxes), so the
                        // results are those of the whole batch on one stream.
                        const int half = (d->P.S + 1) / 2;
                        const int npairs = kc / 2, cur_start = d->cur, bcur_start = casim::box_cur(d);
                        HIPCK(hipEventRecord(d->ev_fork, d->stream));
                        HIPCK(hipStreamWaitEvent(d->stream2, d->ev_fork, 0));
                        for (int h = 0; h < 2; ++h) {
                            d->range_s0 = h ? half : 0;
                            d->range_n = h ? d->P.S - half : half;
                            d->range_stream = h ? d->stream2 : d->stream;
                            d->cur = cur_start;
                            casim::box_set_cur(d, bcur_start);
                            if (h) HIPCK(hipStreamWaitEvent(d->stream2, d->ev_off, 0));
                            for (int i = 0; i < npairs; ++i) {
                                const int ks = 2 * i;  // the pair's first step
                                d->launch_channels(CH_SNAPSHOT | CH_CURRENTS | CH_GATE | (ks ? CH_ACCUMULATE : 0) |
                                                       CH_POOL_UPDATE | CH_POOL_ADVANCE,
                                                   dt);
                                if (!h && !i) HIPCK(hipEventRecord(d->ev_off, d->stream));
                                d->launch_pair(dt);
                                if (d->box) casim::box_efflux_pair(d, ks % M == 0, dt);
                                d->launch_channels(CH_FROM_STORED | CH_CURRENTS | CH_GATE | CH_ACCUMULATE |
                                                       CH_POOL_UPDATE | CH_POOL_ADVANCE,
                                                   dt);
                                if (d->box && (ks + 2) % M == 0) {  // box step on the accumulated effluxes, then the rim
                                    casim::box_step(d, M * dt);
                                    casim::box_rim(d);
                                }
                            }
                        }
                        d->range_s0 = 0;
                        d->range_n = -1;
                        d->range_stream = nullptr;
                        HIPCK(hipEventRecord(d->ev_join, d->stream2));
                        HIPCK(hipStreamWaitEvent(d->stream, d->ev_join, 0));
                        k = 2 * npairs;
                    }
                    for (; k < kc;) {
                        if (pool_rim) {  // the update of the previous step: rim from the cytosol, efflux
                            d->launch_rim(d->Q.cions, d->Q.cbb);
                            d->launch_efflux();
                        }
                        const bool pair = pairs && k + 1 < kc;
                        if (d->box && !pair) casim::box_efflux(d, k % M == 0, dt);
                        // step 1 recomputes the currents of the preceding update, so
                        // it does not add to the accumulated totals again
                        d->launch_channels(CH_SNAPSHOT | CH_CURRENTS | CH_GATE | (k ? CH_ACCUMULATE : 0) |
                                               CH_POOL_UPDATE | CH_POOL_ADVANCE,
                                           dt);
                        if (pair) {
                            // steps k and k+1 in one pass, then step k+1's channel
                            // update from the mouth values the pass stored
                            d->launch_pair(dt);
                            if (d->box) casim::box_efflux_pair(d, k % M == 0, dt);  // both steps' rim effluxes
                            d->launch_channels(CH_FROM_STORED | CH_CURRENTS | CH_GATE | CH_ACCUMULATE | CH_POOL_UPDATE | CH_POOL_ADVANCE,
                                               dt);
                            k += 2;
                        } else {
                            d->launch_step(dt);
                            k += 1;
                        }
                        if (d->box && k % M == 0) {  // box step on the accumulated effluxes, then the rim
                            casim::box_step(d, M * dt);
                            casim::box_rim(d);
                        }
                    }
                    if (pool_rim) d->launch_rim(d->Q.cions, d->Q.cbb);
                    d->launch_efflux();
                    d->launch_channels(CH_SNAPSHOT | CH_CURRENTS | CH_ACCUMULATE | CH_POOL_UPDATE, dt);
                } catch (...) {
                    // rejoin the second stream (if it was forked) so the capture can end
                    if (d->range_n >= 0) {
                        (void)hipEventRecord(d->ev_join, d->stream2);
                        (void)hipStreamWaitEvent(d->stream, d->ev_join, 0);
                    }
                    d->range_s0 = 0;
                    d->range_n = -1;
                    d->range_stream = nullptr;
                    hipGraph_t junk;
                    (void)hipStreamEndCapture(d->stream, &junk);
                    d->cur = cur0;
                    casim::box_set_cur(d, bcur0);
                    throw;
                }
                hipGraph_t graph;
                HIPCK(hipStreamEndCapture(d->stream, &graph));
                hipGraphExec_t exec;
                HIPCK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
                HIPCK(hipGraphDestroy(graph));
                const int flips = (d->cur ^ cur0) | (casim::box_cur(d) ^ bcur0) << 1;
                d->cur = cur0;
                casim::box_set_cur(d, bcur0);
                it = d->graphs.emplace(key, std::make_pair(exec, flips)).first;
            }
            HIPCK(hipGraphLaunch(it->second.first, d->stream));
            d->cur = cur0 ^ (it->second.second & 1);
            casim::box_set_cur(d, bcur0 ^ (it->second.second >> 1));
        }
    });
}

int casim_dyad_step_profiled(CasimDyad *d, double dt, int32_t k_steps, double kernel_ms[2]) {
    return guard([&] {
        REQUIRE(d && kernel_ms, CASIM_E_INVALID, "null argument");
        REQUIRE(k_steps >= 1, CASIM_E_INVALID, "k_steps must be >= 1");
        REQUIRE(d->have_update, CASIM_E_STATE, "step before the first update");
        const int M = casim::box_substeps(d);
        REQUIRE(!d->box || k_steps % M == 0, CASIM_E_INVALID,
                "with a cytosol box, k_steps must be a multiple of its substeps");
        d->use();
        std::vector<hipEvent_t> ev(4 * (size_t)k_steps);
        for (auto &e : ev) HIPCK(hipEventCreate(&e));
        try {
            const bool pool_rim = d->Q.on && d->Q.cyto;
            const bool pairs = d->pair_ok();
            // per step: channel-kernel time in [4k, 4k+1], sweep time in [4k+2, 4k+3]
            // (a two-step pass is timed on its first step, its second channel call
            // on the second step)
            for (int k = 0; k < k_steps;) {
                if (pool_rim) {
                    d->launch_rim(d->Q.cions, d->Q.cbb);
                    d->launch_efflux();
                }
                const bool pair = pairs && k + 1 < k_steps;
                if (d->box && !pair) casim::box_efflux(d, k % M == 0, dt);
                HIPCK(hipEventRecord(ev[4 * k], d->stream));
                d->launch_channels(CH_SNAPSHOT | CH_CURRENTS | CH_GATE | (k ? CH_ACCUMULATE : 0) | CH_POOL_UPDATE |
                                       CH_POOL_ADVANCE,
                                   dt);
                HIPCK(hipEventRecord(ev[4 * k + 1], d->stream));
                HIPCK(hipEventRecord(ev[4 * k + 2], d->stream));
                if (pair) {
                    d->launch_pair(dt);
                    HIPCK(hipEventRecord(ev[4 * k + 3], d->stream));
                    if (d->box) casim::box_efflux_pair(d, k % M == 0, dt);
                    HIPCK(hipEventRecord(ev[4 * (k + 1)], d->stream));
                    d->launch_channels(CH_FROM_STORED | CH_CURRENTS | CH_GATE | CH_ACCUMULATE | CH_POOL_UPDATE | CH_POOL_ADVANCE,
                                           dt);
                    HIPCK(hipEventRecord(ev[4 * (k + 1) + 1], d->stream));
                    HIPCK(hipEventRecord(ev[4 * (k + 1) + 2], d->stream));
                    HIPCK(hipEventRecord(ev[4 * (k + 1) + 3], d->stream));
                    k += 2;
                } else {
                    d->launch_step(dt);
                    HIPCK(hipEventRecord(ev[4 * k + 3], d->stream));
                    k += 1;
                }
                if (d->box && k % M == 0) {
                    casim::box_step(d, M * dt);
                    casim::box_rim(d);
                }
            }
            if (pool_rim) d->launch_rim(d->Q.cions, d->Q.cbb);
            d->launch_efflux();
            d->launch_channels(CH_SNAPSHOT | CH_CURRENTS | CH_ACCUMULATE | CH_POOL_UPDATE, dt);
            HIPCK(hipStreamSynchronize(d->stream));
            double ms_ch = 0, ms_st = 0;
            for (int k = 0; k < k_steps; ++k) {
                float a = 0, b = 0;
                HIPCK(hipEventElapsedTime(&a, ev[4 * k], ev[4 * k + 1]));
                HIPCK(hipEventElapsedTime(&b, ev[4 * k + 2], ev[4 * k + 3]));
                ms_ch += a;
                ms_st += b;
            }
            kernel_ms[0] = ms_st;
            kernel_ms[1] = ms_ch;
        } catch (...) {
            for (auto &e : ev) (void)hipEventDestroy(e);
            throw;
        }
        for (auto &e : ev) (void)hipEventDestroy(e);
    });
}

int casim_dyad_synchronize(CasimDyad *d) {
    return guard([&] {
        REQUIRE(d, CASIM_E_INVALID, "null handle");
        d->use();
        HIPCK(hipStreamSynchronize(d->stream));
    });
}



int casim_dyad_get_effluxes(CasimDyad *d, double *out) {
    return guard([&] {
        REQUIRE(d && out, CASIM_E_INVALID, "null argument");
        d->use();
        copy_out(d, out, d->d_eff, (size_t)(d->P.ni + d->P.nb) * d->P.S);
    });
}

int casim_dyad_get_total_sr_current(CasimDyad *d, double *out) {
    return guard([&] {
        REQUIRE(d && out, CASIM_E_INVALID, "null argument");
        d->use();
        copy_out(d, out, d->d_totals, (size_t)d->P.ni * d->P.S);
    });
}

int casim_dyad_get_release_stats(CasimDyad *d, double *out) {
    return guard([&] {
        REQUIRE(d && out, CASIM_E_INVALID, "null argument");
        d->use();
        std::vector<double> h((size_t)NSTAT * d->P.S);
        copy_out(d, h.data(), d->d_stats, h.size());
        for (int s = 0; s < d->P.S; ++s) {
            const double *st = &h[(size_t)s * NSTAT];
            double *o = out + (size_t)s * 6;
            o[0] = st[ST_AMP];
            o[1] = st[ST_TTP];
            o[2] = st[ST_TTE];
            o[3] = st[ST_Q];
            o[4] = st[ST_TTE] > 0 ? st[ST_Q] / st[ST_TTE] : 0.0;
            o[5] = st[ST_NOPEN];
        }
    });
}

int casim_dyad_get_mouths(CasimDyad *d, double *out) {
    return guard([&] {
        REQUIRE(d && out, CASIM_E_INVALID, "null argument");
        d->use();
        copy_out(d, out, d->d_mouths, (size_t)(d->P.ni + d->P.nb + d->P.nu) * d->P.S * d->M);
    });
}

}  // extern "C"

namespace {
// Locate field fi (ion or bound) in device memory: returns (buffer, slot, is_face).
void field_slot(const CasimDyad *d, int fi, int &slot, bool &face) {
    const Params &P = d->P;
    REQUIRE(fi >= 0 && fi < P.ni + P.nb, CASIM_E_RANGE, "field out of range");
    if (fi < P.ni) {
        slot = fi;
        face = false;
    } else {
        slot = P.bslot[fi - P.ni];
        face = P.bface[fi - P.ni];
    }
}

template <typename T>
void get_field_t(CasimDyad *d, int fi, double *out) {
    const Params &P = d->P;
    int slot;
    bool face;
    field_slot(d, fi, slot, face);
    const size_t N = (size_t)P.S * P.NS;
    if (!face) {
        std::vector<T> h(N);
        HIPCK(hipMemcpyAsync(h.data(), (const T *)d->g[d->cur] + slot * N, sizeof(T) * N, hipMemcpyDeviceToHost,
                             d->stream));
        HIPCK(hipStreamSynchronize(d->stream));
        for (size_t i = 0; i < N; ++i) out[i] = (double)h[i];
    } else {
        const size_t F = (size_t)P.S * 2 * P.nFP;
        std::vector<T> h(F);
        HIPCK(hipMemcpyAsync(h.data(), (const T *)d->f[d->cur] + slot * F, sizeof(T) * F, hipMemcpyDeviceToHost,
                             d->stream));
        HIPCK(hipStreamSynchronize(d->stream));
        std::fill(out, out + N, 0.0);
        for (int s = 0; s < P.S; ++s)
            for (int xy = 0; xy < P.nFP; ++xy) {
                out[(size_t)s * P.NS + (size_t)xy * P.nz] = (double)h[(size_t)s * 2 * P.nFP + xy];
                out[(size_t)s * P.NS + (size_t)xy * P.nz + P.nz - 1] = (double)h[(size_t)s * 2 * P.nFP + P.nFP + xy];
            }
    }
}

template <typename T>
void set_field_t(CasimDyad *d, int fi, const double *in) {
    const Params &P = d->P;
    int slot;
    bool face;
    field_slot(d, fi, slot, face);
    const size_t N = (size_t)P.S * P.NS;
    if (!face) {
        std::vector<T> h(N);
        for (size_t i = 0; i < N; ++i) h[i] = (T)in[i];
        for (int k = 0; k < 2; ++k)
            HIPCK(hipMemcpyAsync((T *)d->g[k] + slot * N, h.data(), sizeof(T) * N, hipMemcpyHostToDevice, d->stream));
        HIPCK(hipStreamSynchronize(d->stream));
    } else {
        const size_t F = (size_t)P.S * 2 * P.nFP;
        std::vector<T> h(F);
        for (int s = 0; s < P.S; ++s)
            for (int xy = 0; xy < P.nFP; ++xy) {
                h[(size_t)s * 2 * P.nFP + xy] = (T)in[(size_t)s * P.NS + (size_t)xy * P.nz];
                h[(size_t)s * 2 * P.nFP + P.nFP + xy] = (T)in[(size_t)s * P.NS + (size_t)xy * P.nz + P.nz - 1];
            }
        for (int k = 0; k < 2; ++k)
            HIPCK(hipMemcpyAsync((T *)d->f[k] + slot * F, h.data(), sizeof(T) * F, hipMemcpyHostToDevice, d->stream));
        HIPCK(hipStreamSynchronize(d->stream));
    }
}
}  // namespace

extern "C" {

int casim_dyad_get_field(CasimDyad *d, int32_t field, double *out) {
    return guard([&] {
        REQUIRE(d && out, CASIM_E_INVALID, "null argument");
        d->use();
        if (d->precision == CASIM_FP64)
            get_field_t<double>(d, field, out);
        else
            get_field_t<float>(d, field, out);
    });
}

int casim_dyad_set_field(CasimDyad *d, int32_t field, const double *in) {
    return guard([&] {
        REQUIRE(d && in, CASIM_E_INVALID, "null argument");
        d->use();
        if (d->precision == CASIM_FP64)
            set_field_t<double>(d, field, in);
        else
            set_field_t<float>(d, field, in);
    });
}

int casim_dyad_get_channel_states(CasimDyad *d, int32_t *out) {
    return guard([&] {
        REQUIRE(d && out, CASIM_E_INVALID, "null argument");
        d->use();
        std::vector<int> h(std::max(1, d->nch));
        HIPCK(hipMemcpyAsync(h.data(), d->d_state, sizeof(int) * d->nch, hipMemcpyDeviceToHost, d->stream));
        HIPCK(hipStreamSynchronize(d->stream));
        for (int s = 0; s < d->P.S; ++s) {
            int o = d->h_off[s], n = d->h_off[s + 1] - o;
            for (int j = 0; j < d->M; ++j) out[(size_t)s * d->M + j] = j < n ? h[o + j] : -1;
        }
    });
}

int casim_dyad_set_channel_states(CasimDyad *d, const int32_t *in) {
    return guard([&] {
        REQUIRE(d && in, CASIM_E_INVALID, "null argument");
        d->use();
        std::vector<int> h(std::max(1, d->nch));
        for (int s = 0; s < d->P.S; ++s) {
            int o = d->h_off[s], n = d->h_off[s + 1] - o;
            for (int j = 0; j < n; ++j) {
                int v = in[(size_t)s * d->M + j];
                REQUIRE(v == 0 || v == 1, CASIM_E_INVALID, "channel state must be 0 or 1");
                h[o + j] = v;
            }
        }
        HIPCK(hipMemcpyAsync(d->d_state, h.data(), sizeof(int) * d->nch, hipMemcpyHostToDevice, d->stream));
        HIPCK(hipStreamSynchronize(d->stream));
    });
}

int casim_dyad_get_rng_states(CasimDyad *d, uint64_t *out) {
    return guard([&] {
        REQUIRE(d && out, CASIM_E_INVALID, "null argument");
        d->use();
        std::vector<uint32_t> h((size_t)d->P.S * 6);
        HIPCK(hipMemcpyAsync(h.data(), d->d_rng, sizeof(uint32_t) * h.size(), hipMemcpyDeviceToHost, d->stream));
        HIPCK(hipStreamSynchronize(d->stream));
        for (size_t i = 0; i < h.size(); ++i) out[i] = h[i];
    });
}

int casim_dyad_set_rng_states(CasimDyad *d, const uint64_t *in) {
    return guard([&] {
        REQUIRE(d && in, CASIM_E_INVALID, "null argument");
        d->use();
        std::vector<uint32_t> h((size_t)d->P.S * 6);
        for (int s = 0; s < d->P.S; ++s) {
            check_seed(in + 6 * (size_t)s);
            for (int i = 0; i < 6; ++i) h[6 * (size_t)s + i] = (uint32_t)in[6 * (size_t)s + i];
        }
        HIPCK(hipMemcpyAsync(d->d_rng, h.data(), sizeof(uint32_t) * h.size(), hipMemcpyHostToDevice, d->stream));
        HIPCK(hipStreamSynchronize(d->stream));
    });
}

double casim_dyad_element_volume(const CasimDyad *d) { return d ? d->elem_volume : 0.0; }

int32_t casim_dyad_max_channels(const CasimDyad *d) { return d ? d->M : 0; }

int32_t casim_dyad_site_channels(const CasimDyad *d, int32_t site) {
    if (!d || site < 0 || site >= d->P.S) return -1;
    return d->h_off[site + 1] - d->h_off[site];
}

int32_t casim_dyad_channel_node(const CasimDyad *d, int32_t site, int32_t ch) {
    if (!d || site < 0 || site >= d->P.S) return -1;
    int o = d->h_off[site], n = d->h_off[site + 1] - o;
    if (ch < 0 || ch >= n) return -1;
    return d->h_node[o + ch];
}

double casim_dyad_bytes_per_site_step(const CasimDyad *d) {
    if (!d) return 0.0;
    const Params &P = d->P;
    const double all = (double)P.NS, inner = (double)(P.nx - 2) * (P.ny - 2) * P.nz;
    const double fall = 2.0 * P.nFP, finner = 2.0 * (P.nx - 2) * (P.ny - 2);
    return (double)d->tsize * (P.n_grid * (all + inner) + P.n_face * (fall + finner));
}

int casim_dyad_set_tuning(CasimDyad *d, int32_t planes_per_block, int32_t threads_per_block, int32_t prefetch) {
    return guard([&] {
        REQUIRE(d, CASIM_E_INVALID, "null handle");
        const Params &P = d->P;
        const int Pl = P.ny * P.nz;
        if (planes_per_block < 0) {  // thread-per-node kernel
            int th = threads_per_block ? threads_per_block : 256;
            REQUIRE(th % 64 == 0 && th >= 64 && th <= 256, CASIM_E_INVALID,
                    "threads must be 64..256, multiple of 64");
            d->threads = th;
            d->use_sweep = false;
        } else {
            REQUIRE(d->sweep_threads > 0, CASIM_E_INVALID, "the plane-marching sweep does not fit this grid");
            int th = threads_per_block ? threads_per_block : d->sweep_threads;
            REQUIRE(th % 64 == 0 && th >= 64 && th <= 512, CASIM_E_INVALID,
                    "threads must be 64..512, multiple of 64");
            REQUIRE((Pl + th - 1) / th <= 4, CASIM_E_INVALID, "too few threads for one plane (at most 4 slots each)");
            d->sweep_threads = th;
            d->sweep_chunk = planes_per_block ? std::min(planes_per_block, P.nx - 2) : P.nx - 2;
            REQUIRE(prefetch >= 0 && prefetch <= 3, CASIM_E_INVALID, "prefetch depth must be 0..3");
            d->sweep_prefetch = prefetch ? prefetch : 1;
            d->use_sweep = true;
        }
        d->clear_graphs();
    });
}

int casim_dyad_record_open_counts(CasimDyad *d, int32_t steps_per_bin, int32_t max_bins) {
    return guard([&] {
        REQUIRE(d, CASIM_E_INVALID, "null handle");
        REQUIRE(steps_per_bin >= 1 && max_bins >= 0, CASIM_E_INVALID, "steps_per_bin >= 1, max_bins >= 0");
        d->use();
        HIPCK(hipStreamSynchronize(d->stream));
        if (d->Ch.trace) (void)hipFree(d->Ch.trace);
        d->Ch.trace = nullptr;
        d->Ch.trace_every = steps_per_bin;
        d->Ch.trace_bins = max_bins;
        if (max_bins > 0) {
            HIPCK(hipMalloc(&d->Ch.trace, (size_t)max_bins * d->P.S));
            HIPCK(hipMemsetAsync(d->Ch.trace, 0xFF, (size_t)max_bins * d->P.S, d->stream));
            HIPCK(hipStreamSynchronize(d->stream));
        }
        d->clear_graphs();
    });
}

int casim_dyad_get_open_counts(CasimDyad *d, uint8_t *out) {
    return guard([&] {
        REQUIRE(d && out, CASIM_E_INVALID, "null argument");
        REQUIRE(d->Ch.trace, CASIM_E_STATE, "open-count recording is off");
        d->use();
        HIPCK(hipMemcpyAsync(out, d->Ch.trace, (size_t)d->Ch.trace_bins * d->P.S, hipMemcpyDeviceToHost,
                             d->stream));
        HIPCK(hipStreamSynchronize(d->stream));
    });
}

int casim_dyad_set_steps_per_pass(CasimDyad *d, int32_t steps) {
    return guard([&] {
        REQUIRE(d, CASIM_E_INVALID, "null handle");
        REQUIRE(steps == 1 || steps == 2, CASIM_E_INVALID, "steps per pass must be 1 or 2");
        d->steps_per_pass = steps;
        d->clear_graphs();
    });
}

int32_t casim_dyad_steps_per_pass(const CasimDyad *d) { return d ? (d->pair_ok() ? 2 : 1) : -1; }

int casim_dyad_set_split(CasimDyad *d, int32_t on) {
    return guard([&] {
        REQUIRE(d, CASIM_E_INVALID, "null handle");
        REQUIRE(on == 0 || on == 1, CASIM_E_INVALID, "on must be 0 or 1");
        d->use();
        HIPCK(hipStreamSynchronize(d->stream));
        if (d->split != (on != 0)) d->clear_graphs();
        d->split = on != 0;
    });
}

void *casim_dyad_stream(CasimDyad *d) { return d ? (void *)d->stream : nullptr; }

}  // extern "C"
