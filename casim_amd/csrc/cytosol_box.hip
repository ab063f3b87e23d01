// cytosol_box.hip -- a 3D cytosol box around every release site, coupled to
// the dyad at its rim.
//
// The reference's coupled solver (DyadRD2DwCytosol.cu:53-59, 119-241) puts
// cleft and cytosol on one 2D grid with two spacings and Dirichlet outer
// edges.  Here each site owns a box of nx*ny*nz nodes at spacing h whose
// z = 0 face carries the dyad; the two grids meet at the dyad's rim:
//   rim      dyad rim node <- bilinear interpolation of the box's z = 0 face
//   efflux   rim link l carries e_l = D (u_in - u_rim) A_l / d_l out of the
//            dyad, the flux the dyad's explicit step removes through it;
//            dt e_l accumulates over the box's substeps
//   deposit  the accumulated amount enters the box nodes with the same
//            bilinear weights (reciprocal to the rim read), divided by the
//            node volume h^3 / 2 of the z = 0 face
// Dyad + box therefore conserve each ion's total exactly in exact arithmetic
// (no channels, SERCA or Dirichlet faces).
//
// Layout in HBM: box fields [field][site][z][y][x] (x fastest), double,
// two buffers; the accumulator [exchanged field][site][link].  The box steps
// every `substeps` dyad steps; its step marches in z with the tile's plane in
// LDS, the rim and efflux kernels are one thread per rim column or link.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "dyad_internal.hpp"

namespace casim {

constexpr int BOX_MAXF = MAXI + MAXB;  // box fields
constexpr int BOX_MAXU = MAXU;         // box buffers

// Everything the box kernels read, passed by value.
struct BoxArgs {
    int nx, ny, nz, NB, S;
    int ni, nbb, nbu, nf;  // box ions, bound fields, buffers, fields (ni + nbb)
    int ne;                // exchanged fields
    int far_dirichlet;
    double ih2, inv_face_volume;
    double D[BOX_MAXF];
    int bion[MAXB], bbuf[MAXB];
    double kon[MAXB], koff[MAXB], total[BOX_MAXU];
    double jmax[MAXI], K2[MAXI], base[MAXI];  // SERCA: Jmax, K^2, c0^2/(c0^2+K^2)
    int ex_box[BOX_MAXF];                     // exchanged field e: box field
    int ex_grid[BOX_MAXF];                    // exchanged field e: dyad grid slot
    double ex_D[BOX_MAXF];                    // exchanged field e: diffusion in the dyad
    // dyad side
    int dnz;            // dyad nodes in z (1 in 2D)
    size_t dfs;         // dyad grid field stride (S * NS)
    int dNS;            // dyad nodes per site
    int ncol, nlink;
    const int *col_node;    // [ncol] dyad flat node of the column at z = 0
    const int *col_box;     // [ncol*4] box face nodes
    const double *col_w;    // [ncol*4] bilinear weights
    const int *link_in;     // [nlink] dyad flat node of the interior end (z = 0)
    const int *link_col;    // [nlink] rim column
    const int *link_ring;   // [nlink] ring index of the interior end (ring_index)
    int RN;                 // ring values per site and grid field (ring nodes x dyad z nodes)
    int ngrid;              // dyad grid fields (ring halves hold every one)
    void *rimc;             // [ne][S][ncol] the rim values k_box_rim wrote (dyad precision)
    const double *link_fac; // [nlink] A / d per unit z weight
    const double *zw;       // [dnz] z weights (1 in 2D; 1/2 on the two z-planes in 3D)
    const int *dep_row;     // [nx*ny] deposit row of a z = 0 node, or -1
    const int *dep_ptr;     // [rows+1]
    const int *dep_link;    // [entries]
    const double *dep_w;    // [entries]
    double *acc;            // [ne][S][nlink]
};

struct CytoBox {
    BoxArgs a{};
    int substeps = 1;
    int cur = 0;
    double *b[2] = {nullptr, nullptr};
    void *ring = nullptr;      // two halves, each [grid field][site][RN], written by k_sweep2: the
                               // pass's step-k+1 values, then its step-k (input) values
    std::vector<double> init;  // [nf] rest values
    double *d_init = nullptr;  // the same on the device
    std::vector<void *> owned;
};

// ---- kernels ---------------------------------------------------------------

__global__ void k_box_init(double *b0, double *b1, const double *__restrict__ init, BoxArgs a) {
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t per = (size_t)a.S * a.NB;
    if (n >= per) return;
    for (int f = 0; f < a.nf; ++f) b0[f * per + n] = b1[f * per + n] = init[f];
}

// Dyad rim nodes from the box: one thread per (rim column, site).
template <typename T>
__global__ void k_box_rim(T *g0, T *g1, const double *__restrict__ box, BoxArgs a, int s0) {
#pragma clang fp contract(off)
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = s0 + blockIdx.y;  // sites s0 .. s0 + gridDim.y - 1
    if (c >= a.ncol) return;
    const size_t per = (size_t)a.S * a.NB;
    const int *q = a.col_box + 4 * c;
    const double *w = a.col_w + 4 * c;
    const size_t node = (size_t)s * a.dNS + a.col_node[c];
    for (int e = 0; e < a.ne; ++e) {
        const double *u = box + a.ex_box[e] * per + (size_t)s * a.NB;
        const double v = w[0] * u[q[0]] + w[1] * u[q[1]] + w[2] * u[q[2]] + w[3] * u[q[3]];
        const size_t o = a.ex_grid[e] * a.dfs + node;
        for (int z = 0; z < a.dnz; ++z) g0[o + z] = g1[o + z] = (T)v;
        ((T *)a.rimc)[((size_t)e * a.S + s) * a.ncol + c] = (T)v;
    }
}

// Rim-link effluxes of the dyad state g, times dt, into the accumulator:
// one thread per (link, site).  After a two-step pass (ring non-null) both
// steps come from the ring the pass wrote next to the rim (contiguous, so no
// strided reads of the interior ends of the y-rim links): step k from its
// second half (the pass's input), step k+1 from its first half; the rim is
// the same for both steps.  The sums are acc + dt e(k), then + dt e(k+1), the
// order of two separate calls.
template <typename T>
__global__ void k_box_efflux(const T *__restrict__ g, const T *__restrict__ ring, BoxArgs a, double dt, int first,
                             int s0) {
#pragma clang fp contract(off)
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = s0 + blockIdx.y;  // sites s0 .. s0 + gridDim.y - 1
    if (l >= a.nlink) return;
    const size_t in = (size_t)s * a.dNS + a.link_in[l];
    const size_t rin = (size_t)s * a.RN + (size_t)a.link_ring[l] * a.dnz;
    const size_t half = (size_t)a.ngrid * a.S * a.RN;
    const double fac = a.link_fac[l];
    const T *rimc = (const T *)a.rimc;
    for (int e = 0; e < a.ne; ++e) {
        // the rim node's value (every z of its column): what k_box_rim wrote there
        const double ur = (double)rimc[((size_t)e * a.S + s) * a.ncol + a.link_col[l]];
        const T *ui = ring ? ring + half + (size_t)a.ex_grid[e] * a.S * a.RN + rin : g + a.ex_grid[e] * a.dfs + in;
        double t = 0.0;
        for (int z = 0; z < a.dnz; ++z) t += a.zw[z] * ((double)ui[z] - ur);
        double *acc = a.acc + ((size_t)e * a.S + s) * a.nlink + l;
        const double ev = a.ex_D[e] * (fac * t);
        double v = first ? dt * ev : *acc + dt * ev;
        if (ring) {
            const T *ur1 = ring + (size_t)a.ex_grid[e] * a.S * a.RN + rin;
            double t1 = 0.0;
            for (int z = 0; z < a.dnz; ++z) t1 += a.zw[z] * ((double)ur1[z] - ur);
            const double ev1 = a.ex_D[e] * (fac * t1);
            v = v + dt * ev1;
        }
        *acc = v;
    }
}

// One explicit box step, marching in z.  A workgroup owns one y-tile of one
// site: rows y0 .. y0+ty-1 of every z-plane plus one halo row on each side.
// Thread t holds node (x, r) = (t % nx, t / nx) of the tile (box row
// y0 - 1 + r) for the whole march, with the z-1, z and z+1 values of every
// field in registers and z+2 loaded while z is computed; the tile's current
// plane goes through LDS (node-major [node][field], double-buffered, one
// LDS-only barrier per plane) for the x and y neighbours.  Every stored value
// is read from HBM once (halo rows a second time, from L2).  Work units are
// dealt so that consecutive tiles of a site land on the same XCD.  The
// arithmetic and its order are those of the 7-point step with mirror
// (no-flux) faces: Laplacian, reactions koff B - kon C (Btot - sum B), SERCA,
// the z = 0 deposit, the Euler update; far_dirichlet holds the faces x = 0,
// nx-1, y = 0, ny-1 and z = nz-1.  Ions and bound fields sit in separate
// arrays whose loops run to the compile-time bounds NI, NB with guards, so
// every array index is a constant and the values stay in registers.
template <int NI, int NB>
__global__ __launch_bounds__(1024) void k_box_step(const double *__restrict__ bin, double *__restrict__ bout, BoxArgs a,
                                                   double dt, int ty, int ntile, int s0, int count) {
#pragma clang fp contract(off)
    constexpr int NT = NI + NB;
    extern __shared__ double blds[];  // [2][tile nodes][NT]
    const int per_xcd = gridDim.x >> 3;
    const int unit = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    const int sl = unit / ntile;  // site within s0 .. s0 + count - 1
    if (sl >= count) return;      // whole workgroup
    const int s = s0 + sl;
    const int y0 = (unit - sl * ntile) * ty;
    const int TN = (ty + 2) * a.nx;
    const int t = threadIdx.x;
    const int x = t % a.nx, r = t / a.nx;
    const int y = y0 - 1 + r;
    const bool live = t < TN && y >= 0 && y < a.ny;
    const bool comp = live && r >= 1 && r <= ty;
    const bool held = a.far_dirichlet && (x == 0 || x == a.nx - 1 || y == 0 || y == a.ny - 1);
    // LDS offsets (in nodes) of the x and y neighbours, mirrored on the faces
    const int oxl = x > 0 ? -1 : 1, oxr = x < a.nx - 1 ? 1 : -1;
    const int oyl = y > 0 ? -a.nx : a.nx, oyr = y < a.ny - 1 ? a.nx : -a.nx;
    const size_t per = (size_t)a.S * a.NB;
    const size_t plane = (size_t)a.nx * a.ny;
    const double *ib = bin + (size_t)s * a.NB;
    double *ob = bout + (size_t)s * a.NB;
    const int nxy = x + a.nx * (live ? y : 0);
    auto on = [&](int f) { return f < NI ? f < a.ni : f - NI < a.nbb; };
    auto fld = [&](int f) { return f < NI ? f : a.ni + f - NI; };
    // Q[k]: plane z - 1 + k of every field (k = 3: the prefetch of z + 2)
    double Q[4][NT];
    auto load = [&](double (&D)[NT], int z) {
#pragma unroll
        for (int f = 0; f < NT; ++f) D[f] = (live && on(f)) ? ib[fld(f) * per + z * plane + nxy] : 0.0;
    };
    load(Q[1], 0);
    if (a.nz > 1) load(Q[2], 1);
    for (int z = 0; z < a.nz; ++z) {
        if (z + 2 < a.nz) load(Q[3], z + 2);
        double *L = blds + (size_t)(z & 1) * TN * NT;
        if (t < TN)
#pragma unroll
            for (int f = 0; f < NT; ++f) L[t * NT + f] = Q[1][f];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
        if (comp) {
            const size_t n = (size_t)nxy + z * plane;
            if (held || (a.far_dirichlet && z == a.nz - 1)) {
#pragma unroll
                for (int f = 0; f < NT; ++f)
                    if (on(f)) ob[fld(f) * per + n] = Q[1][f];
            } else {
                const double *Lc = L + (size_t)t * NT;
                auto diffuse = [&](int f, int fd, double &vo, double &eo) {
                    const double vc = Q[1][f];
                    vo = vc;
                    const double zl = z > 0 ? Q[0][f] : Q[2][f];
                    const double zr = z < a.nz - 1 ? Q[2][f] : Q[0][f];
                    const double lap = (Lc[oxr * NT + f] - 2 * vc + Lc[oxl * NT + f]) * a.ih2 +
                                       (Lc[oyr * NT + f] - 2 * vc + Lc[oyl * NT + f]) * a.ih2 + (zr - 2 * vc + zl) * a.ih2;
                    eo = a.D[fd] * lap;
                };
                double vi[NI], ei[NI], di[NI], vb[NB], eb[NB], db[NB];
#pragma unroll
                for (int j = 0; j < NI; ++j) {
                    vi[j] = ei[j] = di[j] = 0.0;
                    if (j < a.ni) diffuse(j, j, vi[j], ei[j]);
                }
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    vb[b] = eb[b] = db[b] = 0.0;
                    if (b < a.nbb) diffuse(NI + b, a.ni + b, vb[b], eb[b]);
                }
#pragma unroll
                for (int j = 0; j < NI; ++j) {
                    if (j >= a.ni) continue;
                    double R = 0.0;
#pragma unroll
                    for (int b = 0; b < NB; ++b) {
                        if (b >= a.nbb || a.bion[b] != j) continue;
                        // free buffer: total minus its bound fields, in bound-field order
                        double fr = a.total[a.bbuf[b]];
#pragma unroll
                        for (int c = 0; c < NB; ++c)
                            if (c < a.nbb && a.bbuf[c] == a.bbuf[b]) fr -= vb[c];
                        const double R1 = a.koff[b] * vb[b] - a.kon[b] * vi[j] * fr;
                        eb[b] += -R1;
                        R += R1;
                    }
                    ei[j] += R;
                    if (a.jmax[j] != 0.0) {
                        const double c2 = vi[j] * vi[j];
                        ei[j] -= a.jmax[j] * (c2 / (c2 + a.K2[j]) - a.base[j]);
                    }
                }
                if (z == 0) {
                    const int row = a.dep_row[nxy];
                    if (row >= 0) {
                        const int k0 = a.dep_ptr[row], k1 = a.dep_ptr[row + 1];
                        for (int e = 0; e < a.ne; ++e) {
                            const double *acc = a.acc + ((size_t)e * a.S + s) * a.nlink;
                            double sum = 0.0;
                            for (int k = k0; k < k1; ++k) sum += a.dep_w[k] * acc[a.dep_link[k]];
                            const int f = a.ex_box[e];
#pragma unroll
                            for (int j = 0; j < NI; ++j)
                                if (f == j) di[j] = sum;
#pragma unroll
                            for (int b = 0; b < NB; ++b)
                                if (f == a.ni + b) db[b] = sum;
                        }
                    }
                }
#pragma unroll
                for (int j = 0; j < NI; ++j)
                    if (j < a.ni) ob[j * per + n] = vi[j] + ei[j] * dt + di[j] * a.inv_face_volume;
#pragma unroll
                for (int b = 0; b < NB; ++b)
                    if (b < a.nbb) ob[(a.ni + b) * per + n] = vb[b] + eb[b] * dt + db[b] * a.inv_face_volume;
            }
        }
#pragma unroll
        for (int f = 0; f < NT; ++f) {
            Q[0][f] = Q[1][f];
            Q[1][f] = Q[2][f];
            Q[2][f] = Q[3][f];
        }
    }
}

// ---- host side -------------------------------------------------------------

namespace {

template <class T>
T *upload(CytoBox *b, const std::vector<T> &h, hipStream_t st) {
    void *p = nullptr;
    HIPCK(hipMalloc(&p, sizeof(T) * (h.empty() ? 1 : h.size())));
    b->owned.push_back(p);
    if (!h.empty()) {
        HIPCK(hipMemcpyAsync(p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, st));
        HIPCK(hipStreamSynchronize(st));  // h may be a temporary
    }
    return (T *)p;
}

}  // namespace

namespace {
void efflux_launch(CasimDyad *d, const void *g, const void *ring, bool first, double dt) {
    CytoBox *b = d->box;
    dim3 grid((b->a.nlink + 127) / 128, d->range_count());  // the launch range's sites
    if (d->precision == CASIM_FP64)
        hipLaunchKernelGGL(k_box_efflux<double>, grid, dim3(128), 0, d->range_st(), (const double *)g,
                           (const double *)ring, b->a, dt, (int)first, d->range_s0);
    else
        hipLaunchKernelGGL(k_box_efflux<float>, grid, dim3(128), 0, d->range_st(), (const float *)g,
                           (const float *)ring, b->a, dt, (int)first, d->range_s0);
    HIPCK(hipGetLastError());
}
}  // namespace

void box_efflux(CasimDyad *d, bool first, double dt) { efflux_launch(d, d->g[d->cur], nullptr, first, dt); }
// after a two-step pass: the ring holds steps k+1 and k next to the rim (the
// pass's input buffer, now g[cur ^ 1], holds step k everywhere)
void box_efflux_pair(CasimDyad *d, bool first, double dt) {
    efflux_launch(d, d->g[d->cur ^ 1], d->box->ring, first, dt);
}
void *box_ring(const CasimDyad *d) { return d->box ? d->box->ring : nullptr; }

void box_step(CasimDyad *d, double dt) {
    CytoBox *b = d->box;
    const BoxArgs &a = b->a;
    const bool small = a.ni == 1 && a.nbb <= 3;
    const int nt = small ? 4 : MAXI + MAXB;  // fields per LDS node
    // y-tiles of ty rows (+ 2 halo rows): up to 1024 nodes, and two LDS planes
    // within 64 KB; nx <= 64 (attach).  The fewest tiles that fit, with the
    // rows dealt evenly, so that little is re-read as halo (51^3 box: three
    // tiles of 17 rows, 57 rows read per 51, where tiles of 8 rows read 70)
    const int rows = std::min(1024 / a.nx, (int)(65536 / (2 * sizeof(double) * nt * a.nx)));
    const int cap = std::min(a.ny, std::max(1, rows - 2));
    const int ntile = (a.ny + cap - 1) / cap;
    const int ty = (a.ny + ntile - 1) / ntile;
    const int threads = ((ty + 2) * a.nx + 63) / 64 * 64;
    const size_t lds = 2 * sizeof(double) * nt * (ty + 2) * a.nx;
    const long units = (long)ntile * d->range_count();  // the launch range's sites
    REQUIRE(units < (1L << 31) - 8, CASIM_E_INVALID, "cytosol box batch too large for one launch");
    REQUIRE(threads <= 1024 && lds <= 65536, CASIM_E_INVALID, "cytosol box tile does not fit a workgroup");
    dim3 grid((unsigned)((units + 7) / 8 * 8));
    if (small)
        hipLaunchKernelGGL((k_box_step<1, 3>), grid, dim3(threads), lds, d->range_st(), b->b[b->cur],
                           b->b[b->cur ^ 1], a, dt, ty, ntile, d->range_s0, d->range_count());
    else
        hipLaunchKernelGGL((k_box_step<MAXI, MAXB>), grid, dim3(threads), lds, d->range_st(), b->b[b->cur],
                           b->b[b->cur ^ 1], a, dt, ty, ntile, d->range_s0, d->range_count());
    HIPCK(hipGetLastError());
    b->cur ^= 1;
}

void box_rim(CasimDyad *d) {
    CytoBox *b = d->box;
    dim3 grid((b->a.ncol + 127) / 128, d->range_count());  // the launch range's sites
    if (d->precision == CASIM_FP64)
        hipLaunchKernelGGL(k_box_rim<double>, grid, dim3(128), 0, d->range_st(), (double *)d->g[0],
                           (double *)d->g[1], b->b[b->cur], b->a, d->range_s0);
    else
        hipLaunchKernelGGL(k_box_rim<float>, grid, dim3(128), 0, d->range_st(), (float *)d->g[0], (float *)d->g[1],
                           b->b[b->cur], b->a, d->range_s0);
    HIPCK(hipGetLastError());
}

void box_reset(CasimDyad *d) {
    CytoBox *b = d->box;
    const size_t per = (size_t)b->a.S * b->a.NB;
    hipLaunchKernelGGL(k_box_init, dim3((unsigned)((per + 255) / 256)), dim3(256), 0, d->stream, b->b[0], b->b[1],
                       b->d_init, b->a);
    HIPCK(hipGetLastError());
    HIPCK(hipStreamSynchronize(d->stream));
    b->cur = 0;
}

void box_free(CasimDyad *d) {
    if (!d->box) return;
    for (void *p : d->box->owned)
        if (p) (void)hipFree(p);
    delete d->box;
    d->box = nullptr;
}

int box_substeps(const CasimDyad *d) { return d->box ? d->box->substeps : 1; }
int box_cur(const CasimDyad *d) { return d->box ? d->box->cur : 0; }
void box_set_cur(CasimDyad *d, int cur) {
    if (d->box) d->box->cur = cur;
}

}  // namespace casim

using namespace casim;

namespace {

// Bilinear support of the point (px, py) nm on the box's z = 0 face: four
// face nodes and weights, (x0,y0) (x1,y0) (x0,y1) (x1,y1).  A zero weight
// points at the node of the other side so that no face node is referenced
// without need.
void bilinear(const BoxArgs &a, double h, double px, double py, int q[4], double w[4]) {
    const double gx = px / h, gy = py / h;
    const int ax = (int)std::floor(gx), ay = (int)std::floor(gy);
    const double fx = gx - ax, fy = gy - ay;
    REQUIRE(ax >= 1 && ay >= 1 && ax + (fx > 0) <= a.nx - 2 && ay + (fy > 0) <= a.ny - 2, CASIM_E_INVALID,
            "the dyad's rim must lie inside the box's z = 0 face, off its x and y faces");
    const int x1 = fx > 0 ? ax + 1 : ax, y1 = fy > 0 ? ay + 1 : ay;
    q[0] = ax + a.nx * ay;
    q[1] = x1 + a.nx * ay;
    q[2] = ax + a.nx * y1;
    q[3] = x1 + a.nx * y1;
    w[0] = (1 - fx) * (1 - fy);
    w[1] = fx * (1 - fy);
    w[2] = (1 - fx) * fy;
    w[3] = fx * fy;
}

}  // namespace

extern "C" {

int casim_dyad_attach_cytosol_box(CasimDyad *d, const CasimCytosolBoxConfig *c) {
    return guard([&] {
        REQUIRE(d && c, CASIM_E_INVALID, "null argument");
        REQUIRE(!d->box, CASIM_E_STATE, "a cytosol box is already attached");
        REQUIRE(!(d->Q.on && d->Q.cyto), CASIM_E_STATE, "a cytosol pool is attached");
        REQUIRE(c->nx >= 4 && c->ny >= 4 && c->nz >= 2, CASIM_E_INVALID, "box needs nx, ny >= 4 and nz >= 2");
        REQUIRE(c->nx <= 64 && c->nz <= 65535, CASIM_E_INVALID, "box needs nx <= 64 (x-rows of the box step's tiles)");
        REQUIRE((int64_t)c->nx * c->ny * c->nz < (1LL << 30), CASIM_E_INVALID, "box too large");
        REQUIRE(c->h > 0 && std::isfinite(c->h), CASIM_E_INVALID, "box spacing must be > 0");
        REQUIRE(c->substeps >= 1, CASIM_E_INVALID, "substeps must be >= 1");
        REQUIRE(c->ion_D, CASIM_E_INVALID, "null ion_D");
        REQUIRE(c->n_buffers >= 0 && c->n_buffers <= BOX_MAXU, CASIM_E_INVALID, "too many box buffers");
        REQUIRE(c->n_bound >= 0 && c->n_bound <= MAXB, CASIM_E_INVALID, "too many box bound fields");
        REQUIRE(c->n_bound == 0 || (c->buffer_total && c->bound_ion && c->bound_buffer && c->bound_D && c->bound_kon &&
                                    c->bound_koff && c->bound_dyad),
                CASIM_E_INVALID, "null bound-field array");
        const Params &P = d->P;
        d->use();
        auto *b = new CytoBox();
        try {
            BoxArgs &a = b->a;
            a.nx = c->nx;
            a.ny = c->ny;
            a.nz = c->nz;
            a.NB = c->nx * c->ny * c->nz;
            a.S = P.S;
            a.ni = P.ni;
            a.nbb = c->n_bound;
            a.nbu = c->n_buffers;
            a.nf = a.ni + a.nbb;
            a.far_dirichlet = c->far_dirichlet != 0;
            a.ih2 = 1.0 / (c->h * c->h);
            a.inv_face_volume = 2.0 / (c->h * c->h * c->h);
            for (int j = 0; j < P.ni; ++j) {
                REQUIRE(c->ion_D[j] >= 0, CASIM_E_INVALID, "box ion_D must be >= 0");
                a.D[j] = c->ion_D[j];
                const double jm = c->serca_jmax ? c->serca_jmax[j] : 0.0;
                const double K = c->serca_k ? c->serca_k[j] : 0.0;
                REQUIRE(jm == 0.0 || K > 0.0, CASIM_E_INVALID, "SERCA needs K > 0");
                a.jmax[j] = jm;
                a.K2[j] = K * K;
                const double c02 = P.c0[j] * P.c0[j];
                a.base[j] = jm != 0.0 ? c02 / (c02 + a.K2[j]) : 0.0;
            }
            for (int u = 0; u < a.nbu; ++u) a.total[u] = c->buffer_total[u];
            std::vector<int> seen(MAXB, 0);
            for (int k = 0; k < a.nbb; ++k) {
                REQUIRE(c->bound_ion[k] >= 0 && c->bound_ion[k] < P.ni, CASIM_E_INVALID, "bad box bound_ion");
                REQUIRE(c->bound_buffer[k] >= 0 && c->bound_buffer[k] < a.nbu, CASIM_E_INVALID,
                        "bad box bound_buffer");
                REQUIRE(c->bound_D[k] >= 0, CASIM_E_INVALID, "box bound_D must be >= 0");
                a.bion[k] = c->bound_ion[k];
                a.bbuf[k] = c->bound_buffer[k];
                a.D[a.ni + k] = c->bound_D[k];
                a.kon[k] = c->bound_kon[k];
                a.koff[k] = c->bound_koff[k];
                const int db = c->bound_dyad[k];
                REQUIRE(db >= -1 && db < P.nb, CASIM_E_INVALID, "bad box bound_dyad");
                if (db >= 0) {
                    REQUIRE(!seen[db], CASIM_E_INVALID, "two box fields exchange with one dyad field");
                    REQUIRE(P.bion[db] == a.bion[k], CASIM_E_INVALID, "box and dyad bound fields bind different ions");
                    seen[db] = 1;
                }
            }
            // exchanged fields: every ion, then the dyad's mobile bound fields
            a.ne = 0;
            for (int j = 0; j < P.ni; ++j) {
                a.ex_box[a.ne] = j;
                a.ex_grid[a.ne] = j;
                a.ex_D[a.ne] = P.Dion[j];
                ++a.ne;
            }
            for (int db = 0; db < P.nb; ++db) {
                int k = -1;
                for (int t = 0; t < a.nbb; ++t)
                    if (c->bound_dyad[t] == db) k = t;
                if (P.bface[db] || P.Db[db] == 0.0) continue;  // immobile: its rim value is never read
                REQUIRE(k >= 0, CASIM_E_INVALID, "a mobile dyad bound field has no box field to exchange with");
                a.ex_box[a.ne] = a.ni + k;
                a.ex_grid[a.ne] = P.bslot[db];
                a.ex_D[a.ne] = P.Db[db];
                ++a.ne;
            }
            // rest state: resting ions, buffers at equilibrium over the ions they bind
            b->init.assign(a.nf, 0.0);
            for (int j = 0; j < P.ni; ++j) b->init[j] = P.c0[j];
            if (d->start_equilibrium) {
                for (int u = 0; u < a.nbu; ++u) {
                    std::vector<double> kon, koff, cc;
                    std::vector<int> idx;
                    for (int k = 0; k < a.nbb; ++k)
                        if (a.bbuf[k] == u) {
                            kon.push_back(a.kon[k]);
                            koff.push_back(a.koff[k]);
                            cc.push_back(P.c0[a.bion[k]]);
                            idx.push_back(k);
                        }
                    if (idx.empty()) continue;
                    auto sol = buffer_equilibrium(kon, koff, cc, a.total[u]);
                    for (size_t t = 0; t < idx.size(); ++t) b->init[a.ni + idx[t]] = sol[t];
                }
            }
            // dyad rim columns and links
            const int nx = P.nx, ny = P.ny, nzd = P.nz;
            a.dnz = P.dim == 3 ? nzd : 1;
            a.dfs = (size_t)P.S * P.NS;
            a.dNS = P.NS;
            a.ncol = 2 * nx + 2 * (ny - 2);
            a.nlink = 2 * (nx - 2) + 2 * (ny - 2);
            std::vector<int> col_node(a.ncol), col_box(4 * a.ncol);
            std::vector<double> col_w(4 * a.ncol);
            auto column = [&](int cidx, int &x, int &y) {
                if (cidx < nx) {
                    x = cidx, y = 0;
                } else if (cidx < 2 * nx) {
                    x = cidx - nx, y = ny - 1;
                } else if (cidx < 2 * nx + ny - 2) {
                    x = 0, y = cidx - 2 * nx + 1;
                } else {
                    x = nx - 1, y = cidx - (2 * nx + ny - 2) + 1;
                }
            };
            auto colof = [&](int x, int y) {
                if (y == 0) return x;
                if (y == ny - 1) return nx + x;
                if (x == 0) return 2 * nx + y - 1;
                return 2 * nx + ny - 2 + y - 1;
            };
            for (int k = 0; k < a.ncol; ++k) {
                int x, y;
                column(k, x, y);
                col_node[k] = (x * ny + y) * nzd;
                bilinear(a, c->h, c->origin_x + x * d->dx, c->origin_y + y * d->dy, &col_box[4 * k], &col_w[4 * k]);
            }
            std::vector<int> link_in(a.nlink), link_col(a.nlink), link_ring(a.nlink);
            std::vector<double> link_fac(a.nlink);
            const double thick = P.dim == 3 ? d->dz : d->zext;
            for (int l = 0; l < a.nlink; ++l) {
                int x, y, xi, yi;
                double fac;
                if (l < 2 * (ny - 2)) {
                    const int side = l >= ny - 2;
                    y = 1 + l - side * (ny - 2);
                    x = side ? nx - 1 : 0;
                    xi = side ? nx - 2 : 1;
                    yi = y;
                    fac = d->dy * thick / d->dx;
                } else {
                    const int r = l - 2 * (ny - 2);
                    const int side = r >= nx - 2;
                    x = 1 + r - side * (nx - 2);
                    y = side ? ny - 1 : 0;
                    yi = side ? ny - 2 : 1;
                    xi = x;
                    fac = d->dx * thick / d->dy;
                }
                link_in[l] = (xi * ny + yi) * nzd;
                link_ring[l] = ring_index(xi, yi, nx, ny);
                link_col[l] = colof(x, y);
                link_fac[l] = fac;
            }
            std::vector<double> zw(a.dnz, 1.0);
            if (P.dim == 3) zw[0] = zw[a.dnz - 1] = 0.5;
            // deposit rows: face nodes in link order
            std::vector<std::vector<std::pair<int, double>>> rows(a.nx * a.ny);
            for (int l = 0; l < a.nlink; ++l) {
                const int k = link_col[l];
                for (int q = 0; q < 4; ++q)
                    if (col_w[4 * k + q] != 0.0) rows[col_box[4 * k + q]].push_back({l, col_w[4 * k + q]});
            }
            std::vector<int> dep_row(a.nx * a.ny, -1), dep_ptr(1, 0), dep_link;
            std::vector<double> dep_w;
            for (int n = 0; n < a.nx * a.ny; ++n) {
                if (rows[n].empty()) continue;
                dep_row[n] = (int)dep_ptr.size() - 1;
                for (auto &e : rows[n]) {
                    dep_link.push_back(e.first);
                    dep_w.push_back(e.second);
                }
                dep_ptr.push_back((int)dep_link.size());
            }
            a.col_node = upload(b, col_node, d->stream);
            a.col_box = upload(b, col_box, d->stream);
            a.col_w = upload(b, col_w, d->stream);
            a.link_in = upload(b, link_in, d->stream);
            a.link_col = upload(b, link_col, d->stream);
            a.link_ring = upload(b, link_ring, d->stream);
            a.RN = ring_nodes(nx, ny) * a.dnz;
            a.ngrid = P.n_grid;
            HIPCK(hipMalloc(&b->ring, 2 * d->tsize * (size_t)P.n_grid * P.S * a.RN));
            b->owned.push_back(b->ring);
            HIPCK(hipMalloc(&a.rimc, d->tsize * (size_t)a.ne * P.S * a.ncol));
            b->owned.push_back(a.rimc);
            a.link_fac = upload(b, link_fac, d->stream);
            a.zw = upload(b, zw, d->stream);
            a.dep_row = upload(b, dep_row, d->stream);
            a.dep_ptr = upload(b, dep_ptr, d->stream);
            a.dep_link = upload(b, dep_link, d->stream);
            a.dep_w = upload(b, dep_w, d->stream);
            a.acc = upload(b, std::vector<double>((size_t)a.ne * P.S * a.nlink, 0.0), d->stream);
            const size_t per = (size_t)a.nf * P.S * a.NB;
            for (int i = 0; i < 2; ++i) {
                HIPCK(hipMalloc(&b->b[i], sizeof(double) * per));
                b->owned.push_back(b->b[i]);
            }
            b->d_init = upload(b, b->init, d->stream);
            b->substeps = c->substeps;
            d->box = b;
            box_reset(d);
        } catch (...) {
            if (d->box == b) d->box = nullptr;
            for (void *p : b->owned)
                if (p) (void)hipFree(p);
            delete b;
            throw;
        }
        d->clear_graphs();
        d->have_update = false;
    });
}

int32_t casim_dyad_cytosol_box_fields(const CasimDyad *d) { return d && d->box ? d->box->a.nf : 0; }

int casim_dyad_get_cytosol_box(CasimDyad *d, int32_t field, double *out) {
    return guard([&] {
        REQUIRE(d && out, CASIM_E_INVALID, "null argument");
        REQUIRE(d->box, CASIM_E_STATE, "no cytosol box attached");
        REQUIRE(field >= 0 && field < d->box->a.nf, CASIM_E_RANGE, "box field out of range");
        d->use();
        const size_t per = (size_t)d->P.S * d->box->a.NB;
        HIPCK(hipMemcpyAsync(out, d->box->b[d->box->cur] + field * per, sizeof(double) * per, hipMemcpyDeviceToHost,
                             d->stream));
        HIPCK(hipStreamSynchronize(d->stream));
    });
}

int casim_dyad_set_cytosol_box(CasimDyad *d, int32_t field, const double *in) {
    return guard([&] {
        REQUIRE(d && in, CASIM_E_INVALID, "null argument");
        REQUIRE(d->box, CASIM_E_STATE, "no cytosol box attached");
        REQUIRE(field >= 0 && field < d->box->a.nf, CASIM_E_RANGE, "box field out of range");
        d->use();
        const size_t per = (size_t)d->P.S * d->box->a.NB;
        for (int i = 0; i < 2; ++i)
            HIPCK(hipMemcpyAsync(d->box->b[i] + field * per, in, sizeof(double) * per, hipMemcpyHostToDevice,
                                 d->stream));
        HIPCK(hipStreamSynchronize(d->stream));
    });
}

}  // extern "C"
