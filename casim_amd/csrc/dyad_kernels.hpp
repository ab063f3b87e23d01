// dyad_kernels.hpp -- gfx950 kernels of the dyad reaction-diffusion step.
//
// Device layout (per handle, two ping-pong copies of each):
//   grid fields  [g][site][x][y][z]   g = ions, then bound fields that exist
//                                     on every plane (z fastest, as the
//                                     reference's IndexMP / _Index)
//   face fields  [f][site][face][x][y] bound fields of membrane-only buffers,
//                                     face 0 = plane z = 0, face 1 = z = nz-1
// Free buffer = total - sum of its bound fields is derived where needed, not
// stored.  Rim nodes keep their values through a step: the plane sweeps write
// them back unchanged (so each plane's stores cover whole cache lines), the
// thread-per-node kernel leaves them, and the rim kernels write both copies,
// so both copies carry the same rim.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mrg32k3a.hpp"

namespace casim {

constexpr int MAXI = 4;   // ions
constexpr int MAXB = 8;   // bound fields
constexpr int MAXU = 8;   // unique buffers
constexpr int MAXT = 4;   // channel types
constexpr int NPAR = 16;  // kinetic parameters per type
constexpr int MAX_SITE_CHANNELS = 127;
constexpr int JUMP_N = 256;  // jump table entries: A^(2p), p < JUMP_N

struct Params {
    int dim, nx, ny, nz, S;
    int NS;    // nodes per site
    int nFP;   // nodes per face plane (nx*ny)
    int ni, nb, nu;
    int n_grid, n_face;
    int bion[MAXB], bbuf[MAXB], bface[MAXB], bslot[MAXB];  // slot: grid index or face index
    double Dion[MAXI], Db[MAXB], kon[MAXB], koff[MAXB], binit[MAXB];
    double total[MAXU];
    int ufaces[MAXU];
    int sl[MAXI];
    double N1[MAXI], K1[MAXI], N2[MAXI], K2[MAXI];
    double c0[MAXI];
    double gD[MAXI + MAXB];  // diffusion coefficient of grid field g
    double ihx2, ihy2, ihz2;
    double cc;  // 2 (ihx2 + ihy2 + ihz2): the centre weight of the Laplacian
};

struct Channels {
    const int *off;     // [S+1]
    const int *node;    // [nch] flat node of channel (channel order)
    const int *type;    // [nch]
    const int *perm;    // [nch] position of the channel in its site's node-sorted list
    const int *snode;   // [nch] node-sorted list of mouth nodes per site
    const int *sperm;   // [nch] channel (index within its site) at each node-sorted position
    int *state;         // [nch]
    uint32_t *rng;      // [S*6]
    const uint32_t *jump;  // [JUMP_N*18]
    double *mouths;     // [K*S*M]  K = ni + nb + nu
    double *snap;       // [(n_grid + n_face) * nch] step-k+1 values of the mouth nodes, node-sorted,
                        // stored by k_sweep2 (face values 0 off the two z-faces)
    double *pg;         // [ni * nch] node-sorted: state * conductance after the last gating
    double *pjv;        // [ni * nch] node-sorted: the jSR each channel sees (constant jSR, no pools)
    const double *jsr;  // [M*S*ni]
    double *sJ;         // [ni*nch] currents in node-sorted order
    double *totals;     // [ni*S]
    double *stats;      // [S*NSTAT] release statistics, see StatSlot
    uint8_t *trace;     // [bins][S] open channels sampled every trace_every steps (or null)
    int trace_every, trace_bins;
    int M;              // max channels per site
    int nch;            // channels over all sites
    int ca, mg;
    int model[MAXT];
    double par[MAXT * NPAR];
    double g[MAXT * MAXI];
};

enum ChannelMode {
    CH_SNAPSHOT = 1,
    CH_CURRENTS = 2,
    CH_GATE = 4,
    CH_GATE_STORED = 8,
    CH_ACCUMULATE = 16,
    CH_POOL_UPDATE = 32,   // SRPool::Update from the fresh site totals
    CH_POOL_ADVANCE = 64,  // CytosolPool::RunRD and SRPool::Run for the next step
    CH_FROM_STORED = 128,  // take the mouth values from mouths[] (written by k_sweep2) instead of the grid
};

constexpr int MAXCU = 12;  // cytosol buffers (dyad buffers first)
constexpr int MAXCB = 16;  // cytosol bound fields

// Compartments around the dyad held on the device (SRPool.cpp,
// NSRConstPool.cpp, CytosolPool.cpp), one value per site and species.
struct Pools {
    int on;           // pools attached
    int jsr_const, cyto;
    double cq, Vjsr, Velem;
    double nCQ[MAXI], KCQ[MAXI], tau[MAXI];
    double *jions, *jnew, *IjSR, *jnsr;  // [site + ion*S]
    const double *nsr;                   // [site + ion*S]
    int nbc, nBc;
    int cbind[MAXI * MAXCU];             // [ion * nbc + buffer] -> bound index or -1
    int cb_dyad[MAXCB];
    double ctot[MAXCU], ckon[MAXCB], ckoff[MAXCB], conet_b[MAXCB], cinit[MAXCB];
    double conet_i[MAXI], cbg[MAXI];
    int serca[MAXI];
    double jmax[MAXI], kup[MAXI];
    double *cions, *cbuf, *cfree, *cbb;  // [site + k*S]
    const double *eff;                   // dyad effluxes [f + (ni+nb)*site]
};

// SRPool::Update (SRPool.cpp:55-68) for one site.
static __device__ void pool_update_site(const Pools &q, int s, int S, int ni, const double *totals) {
#pragma clang fp contract(off)
    if (q.jsr_const) return;
    for (int i = 0; i < ni; ++i) {
        q.IjSR[s + i * S] = totals[s + i * S] * q.Velem;
        q.jnsr[s + i * S] = q.nsr[s + i * S];
        q.jions[s + i * S] = q.jnew[s + i * S];
    }
}

// CytosolPool::RunRD (CytosolPool.cpp:217-268) and SRPool::Run
// (SRPool.cpp:70-89) for one site, in the reference's operation order.
static __device__ void pool_advance_site(const Pools &q, int s, int S, int ni, int nb, double dt) {
#pragma clang fp contract(off)
    if (q.cyto) {
        const int nIB = ni + nb;
        double evo_i[MAXI], evo_b[MAXCB];
        int kd = 0, ks = 0;
        for (int j = 0; j < ni; ++j) {
            double react = 0;
            for (int u = 0; u < q.nbc; ++u) {
                const int b = q.cbind[j * q.nbc + u];
                if (b < 0) continue;
                const double R1 = q.ckoff[b] * q.cbuf[s + b * S] - q.ckon[b] * q.cions[s + j * S] * q.cfree[s + u * S];
                evo_b[b] = -R1;
                react += R1;
                if (q.cb_dyad[b]) {
                    evo_b[b] += q.eff[ni + kd + s * nIB];
                    ++kd;
                }
            }
            evo_i[j] = react + q.eff[j + s * nIB];
            if (q.serca[j]) {
                const double c = q.cions[s + j * S];
                const double ions2 = c * c;
                const double K2 = q.kup[j] * q.kup[j];
                const double b2 = q.cbg[ks] * q.cbg[ks];  // indexed by the SERCA counter, as the reference
                evo_i[j] -= q.jmax[j] * (ions2 / (ions2 + K2) - b2 / (b2 + K2));
                ++ks;
            }
        }
        for (int j = 0; j < ni; ++j) {
            const double c = q.cions[s + j * S];
            evo_i[j] -= q.conet_i[j] * (c - q.cbg[j]);
            q.cions[s + j * S] = c + evo_i[j] * dt;
        }
        kd = 0;
        for (int b = 0; b < q.nBc; ++b) {
            const double c = q.cbuf[s + b * S];
            evo_b[b] -= q.conet_b[b] * (c - q.cinit[b]);
            const double nv = c + evo_b[b] * dt;
            q.cbuf[s + b * S] = nv;
            if (q.cb_dyad[b]) q.cbb[s + (kd++) * S] = nv;
        }
        for (int j = 0; j < ni; ++j) {
            if (j == 0)
                for (int u = 0; u < q.nbc; ++u) q.cfree[s + u * S] = q.ctot[u];
            for (int u = 0; u < q.nbc; ++u) {
                const int b = q.cbind[j * q.nbc + u];
                if (b >= 0) q.cfree[s + u * S] -= q.cbuf[s + b * S];
            }
        }
    }
    if (!q.jsr_const) {
        double beta = 1;
        for (int j = 0; j < ni; ++j) {
            const double c = q.jions[s + j * S];
            beta += q.nCQ[j] * q.KCQ[j] * q.cq / ((q.KCQ[j] + c) * (q.KCQ[j] + c));
        }
        for (int i = 0; i < ni; ++i) {
            const double c = q.jions[s + i * S];
            const double Iref = 1 / q.tau[i] * (q.jnsr[s + i * S] - c);
            q.jnew[s + i * S] = c + dt / beta * (Iref - q.IjSR[s + i * S] / q.Vjsr);
        }
    }
}

// Per-site release statistics, the columns of IndChannelCREAnalyzer's
// scenarios table (IndChannelCREAnalyzer.cpp:89-137) accumulated step by step:
// N_j = open channels after the gating of step j, in effect from t_j = j dt.
enum StatSlot {
    ST_STEPS = 0,   // steps gated since InitOpening
    ST_NOPEN = 1,   // N of the last step
    ST_AMP = 2,     // max N
    ST_TTP = 3,     // first time N reached its max
    ST_TTE = 4,     // time of the last change of N
    ST_Q = 5,       // sum of N dt
    NSTAT = 8
};

template <typename T>
__device__ __forceinline__ T load_face(const T *f, const Params &p, int slot, int s, int face, int xy) {
    return f[((size_t)slot * p.S + s) * 2 * p.nFP + (size_t)face * p.nFP + xy];
}

// Total of unique buffer u at plane z: membrane-only buffers vanish off
// their two planes.
__device__ __forceinline__ double total_at(const Params &p, int u, int z) {
    return (p.ufaces[u] && p.dim == 3 && z != 0 && z != p.nz - 1) ? 0.0 : p.total[u];
}

// ---------------------------------------------------------------------------
// The step: for every interior node of every site, the 7-point (5-point)
// Laplacian of each diffusing field, the buffer reactions, the RyR current at
// mouth nodes, the sarcolemmal rapid-buffer factor, and the explicit Euler
// update (DyadRDCPU.cpp:404-449 _EvolutionOp, :472-492 _UpdateEvo, :451-470
// _GetFreeBuffers, with RunRD's Jacobi order :495-506).
// One thread per node; one block row of blocks per site.
// NI/NB > 0: compile-time species counts; 0: runtime counts up to MAXI/MAXB.
template <typename T, int DIM, int NI, int NB>
__global__ __launch_bounds__(256) void k_step(const T *__restrict__ gin, T *__restrict__ gout,
                                              const T *__restrict__ fin, T *__restrict__ fout,
                                              const int *__restrict__ ch_off, const int *__restrict__ snode,
                                              const double *__restrict__ sJ, int nch, Params p, double dt_) {
    constexpr int LI = NI ? NI : MAXI;
    constexpr int LB = NB ? NB : MAXB;
    const int s = blockIdx.y;
    const int n0 = blockIdx.x * blockDim.x;
    const int n = n0 + threadIdx.x;
    const T dt = (T)dt_;

    // Mouth entries of this block: node-sorted list, binary search (uniform).
    int lo = __builtin_amdgcn_readfirstlane(ch_off[s]);
    int hi = __builtin_amdgcn_readfirstlane(ch_off[s + 1]);
    {
        int a = lo, b = hi;
        while (a < b) {
            int m = (a + b) >> 1;
            if (snode[m] < n0) a = m + 1; else b = m;
        }
        int c = a, e = hi;
        const int n1 = n0 + (int)blockDim.x;
        while (c < e) {
            int m = (c + e) >> 1;
            if (snode[m] < n1) c = m + 1; else e = m;
        }
        lo = a;
        hi = c;
    }

    if (n >= p.NS) return;
    const int z = DIM == 3 ? n % p.nz : 0;
    const int t = DIM == 3 ? n / p.nz : n;
    const int y = t % p.ny;
    const int x = t / p.ny;
    if (x == 0 || x == p.nx - 1 || y == 0 || y == p.ny - 1) return;

    const size_t fs = (size_t)p.S * p.NS;  // field stride
    const size_t c = (size_t)s * p.NS + n;
    const int sx = p.ny * p.nz, sy = p.nz;
    const int zl = DIM == 3 ? (z == 0 ? 1 : -1) : 0;
    const int zr = DIM == 3 ? (z == p.nz - 1 ? -1 : 1) : 0;
    const T ihx2 = (T)p.ihx2, ihy2 = (T)p.ihy2, ihz2 = (T)p.ihz2;

    T C[LI], evoC[LI], B[LB], evoB[LB];
    const bool face = DIM == 3 && (z == 0 || z == p.nz - 1);
    const int xy = x * p.ny + y;
    const T cc = (T)p.cc;

    // the arithmetic of node_update (neighbour pairs summed first, the
    // diffusion coefficient applied inside the evolution)
#pragma unroll
    for (int j = 0; j < LI; ++j) {
        if (!NI && j >= p.ni) break;
        const T *u = gin + j * fs + c;
        T vc = u[0];
        C[j] = vc;
        T lap = (u[sx] + u[-sx]) * ihx2 + (u[sy] + u[-sy]) * ihy2;
        if (DIM == 3) lap = lap + (u[zr] + u[zl]) * ihz2;
        evoC[j] = lap - cc * vc;  // the Laplacian until the evolution below
    }
    T lB[LB];
#pragma unroll
    for (int b = 0; b < LB; ++b) {
        if (!NB && b >= p.nb) break;
        lB[b] = 0;
        if (p.bface[b]) {
            B[b] = face ? load_face(fin, p, p.bslot[b], s, z == 0 ? 0 : 1, xy) : (T)0;
        } else {
            const T *u = gin + p.bslot[b] * fs + c;
            T vc = u[0];
            B[b] = vc;
            if (p.Db[b] != 0.0) {
                T lap = (u[sx] + u[-sx]) * ihx2 + (u[sy] + u[-sy]) * ihy2;
                if (DIM == 3) lap = lap + (u[zr] + u[zl]) * ihz2;
                lB[b] = lap - cc * vc;
            }
        }
    }
    // free buffer seen by each bound field, subtracted in bound order
    T frb[LB];
#pragma unroll
    for (int b = 0; b < LB; ++b) {
        if (!NB && b >= p.nb) break;
        T v = (T)total_at(p, p.bbuf[b], z);
#pragma unroll
        for (int b2 = 0; b2 < LB; ++b2) {
            if (!NB && b2 >= p.nb) break;
            if (p.bbuf[b2] == p.bbuf[b]) v -= B[b2];
        }
        frb[b] = v;
    }
    // reactions R1 = koff B - kon C Bfree
    T R[LI];
#pragma unroll
    for (int j = 0; j < LI; ++j) R[j] = (T)-0.0;
#pragma unroll
    for (int b = 0; b < LB; ++b) {
        if (!NB && b >= p.nb) break;
        const int j = p.bion[b];
        T cj = (T)0;
#pragma unroll
        for (int jj = 0; jj < LI; ++jj)
            if (jj == j) cj = C[jj];
        const T R1 = (T)p.koff[b] * B[b] - (T)p.kon[b] * cj * frb[b];
        evoB[b] = p.Db[b] != 0.0 && !p.bface[b] ? (T)p.Db[b] * lB[b] - R1 : -R1;
        const bool on = !p.bface[b] || face;
#pragma unroll
        for (int jj = 0; jj < LI; ++jj)
            if (jj == j && on) R[jj] += R1;
    }
    // currents at mouth nodes (the last channel on a node wins, as the
    // reference's assignment does)
    T Jn[LI];
#pragma unroll
    for (int j = 0; j < LI; ++j) Jn[j] = 0;
    for (int k = lo; k < hi; ++k) {
        if (snode[k] == n) {
#pragma unroll
            for (int j = 0; j < LI; ++j) {
                if (!NI && j >= p.ni) break;
                Jn[j] = (T)sJ[(size_t)j * nch + k];
            }
        }
    }
#pragma unroll
    for (int j = 0; j < LI; ++j) {
        if (!NI && j >= p.ni) break;
        T e = (T)p.Dion[j] * evoC[j] + (R[j] + Jn[j]);
        if ((DIM == 2 || z == p.nz - 1) && p.sl[j]) {
            T s1 = (T)p.K1[j] + C[j];
            T s2 = (T)p.K2[j] + C[j];
            e *= (T)1 / ((T)1 + (T)p.N1[j] * (T)p.K1[j] / (s1 * s1) + (T)p.N2[j] * (T)p.K2[j] / (s2 * s2));
        }
        gout[j * fs + c] = C[j] + e * dt;
    }
#pragma unroll
    for (int b = 0; b < LB; ++b) {
        if (!NB && b >= p.nb) break;
        if (p.bface[b]) {
            if (face)
                fout[((size_t)p.bslot[b] * p.S + s) * 2 * p.nFP + (size_t)(z == 0 ? 0 : 1) * p.nFP + xy] =
                    B[b] + evoB[b] * dt;
        } else {
            gout[p.bslot[b] * fs + c] = B[b] + evoB[b] * dt;
        }
    }
}


// Workgroup barrier that orders LDS traffic only: global loads issued before
// it stay in flight (the plane prefetch below relies on that).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Element at a 32-bit byte offset from a wave-uniform base: the address is
// base + zext(offset), which the compiler issues as a global access with the
// base in scalar registers and one 32-bit vector offset, instead of a 64-bit
// address per lane and field.  Offsets stay below 4 GiB because they never
// leave one site's field (NS nodes).
template <typename T>
__device__ __forceinline__ T ld_off(const T *base, unsigned boff) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + boff);
}
template <typename T>
__device__ __forceinline__ void st_off(T *base, unsigned boff, T v) {
    *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + boff) = v;
}

// Species layouts the sweeps can take at compile time: bound field b binds
// ion bion[b], belongs to buffer bbuf[b], lives on the z-faces only when
// bface[b] and sits in grid (or face) slot bslot[b]; ufaces[u]: buffer u
// exists on the two z-faces only; the fixed layouts have no sarcolemmal
// buffer and diffusing A and B.  LY = 0 reads all of it from Params.
//  1: one ion; buffers A, B mobile, F on the z-faces (Ca-A, Ca-B, Ca-F)
//  2: one ion; buffers A, B mobile, F fixed on every node
template <int LY>
struct Layout {
    static constexpr bool fixed = false;
    static constexpr int nu = MAXU;
    static constexpr int bion[1] = {0}, bbuf[1] = {0}, bface[1] = {0}, bslot[1] = {0}, ufaces[1] = {0};
};
template <>
struct Layout<1> {
    static constexpr bool fixed = true;
    static constexpr int nu = 3;
    static constexpr int bion[3] = {0, 0, 0}, bbuf[3] = {0, 1, 2}, bface[3] = {0, 0, 1}, bslot[3] = {1, 2, 0};
    static constexpr int ufaces[3] = {0, 0, 1};
};
template <>
struct Layout<2> {
    static constexpr bool fixed = true;
    static constexpr int nu = 3;
    static constexpr int bion[3] = {0, 0, 0}, bbuf[3] = {0, 1, 2}, bface[3] = {0, 0, 0}, bslot[3] = {1, 2, 3};
    static constexpr int ufaces[3] = {0, 0, 0};
};
template <int LY> __device__ __forceinline__ int ly_bion(const Params &p, int b) {
    if constexpr (Layout<LY>::fixed) return Layout<LY>::bion[b]; else return p.bion[b];
}
template <int LY> __device__ __forceinline__ int ly_bbuf(const Params &p, int b) {
    if constexpr (Layout<LY>::fixed) return Layout<LY>::bbuf[b]; else return p.bbuf[b];
}
template <int LY> __device__ __forceinline__ int ly_bface(const Params &p, int b) {
    if constexpr (Layout<LY>::fixed) return Layout<LY>::bface[b]; else return p.bface[b];
}
template <int LY> __device__ __forceinline__ int ly_bslot(const Params &p, int b) {
    if constexpr (Layout<LY>::fixed) return Layout<LY>::bslot[b]; else return p.bslot[b];
}
// sarcolemmal buffer of ion j present (Params::sl)
template <int LY> __device__ __forceinline__ bool ly_sl(const Params &p, int j) {
    if constexpr (Layout<LY>::fixed) return false; else return p.sl[j] != 0;
}
// bound field b diffuses (a grid field with D != 0)
template <int LY> __device__ __forceinline__ bool ly_mobile(const Params &p, int b) {
    if constexpr (Layout<LY>::fixed) return !Layout<LY>::bface[b] && b < 2; else return p.Db[b] != 0.0;
}
template <int LY> __device__ __forceinline__ double ly_total(const Params &p, int u, int z) {
    int faces;
    if constexpr (Layout<LY>::fixed) faces = Layout<LY>::ufaces[u]; else faces = p.ufaces[u];
    return (faces && p.dim == 3 && z != 0 && z != p.nz - 1) ? 0.0 : p.total[u];
}

// Updated values of one interior node (DyadRDCPU.cpp:404-492): the
// Laplacian of every diffusing field (x-neighbours from registers, y and z
// from the LDS plane), the reactions with the free buffers seen by each
// bound field, the channel current, the sarcolemmal factor at z = nz-1 (3D)
// or everywhere (2D), the Euler step.  C0/Cm/Cp: the node and its x-1 / x+1
// neighbours, every grid field; Fv: face values of the node.  The LDS plane
// is node-major ([node][grid field]): Lc points at the node's first field,
// its y-neighbours sit sy elements away and its z-neighbours zl, zr elements
// away (the mirror image on the z-faces).  Writes the new grid values to G
// (every grid field) and, at face nodes, the new face values to Fo.
//
// The reference's terms, regrouped for FP64 issue: each Laplacian sums the
// two neighbours of a direction first, ihx2 (u[x+1] + u[x-1]) + ihy2 (..)
// + ihz2 (..) - cc u with cc = 2 (ihx2 + ihy2 + ihz2), and the diffusion
// coefficient multiplies it inside the evolution, D lap + (reactions +
// current) for an ion and D lap - R1 for a bound field (7 + 2 operations
// per diffusing field where the reference's order takes 10 + 3).  The
// values differ from the reference's order by rounding only.
template <typename T, int DIM, int NI, int NB, int NG, int NF, int LY, int LI_ = NI ? NI : MAXI>
__device__ __forceinline__ void node_update(const Params &p, const T *Cm, const T (&C0)[NG],
                                            const T (&Cp)[NG], const T *Lc, int sy, int zl, int zr, int z,
                                            bool face, const T (&Fv)[NF], const T (&J)[LI_], T dt, T ihx2, T ihy2, T ihz2,
                                            T (&G)[NG], T (&Fo)[NF]) {
    constexpr int LI = NI ? NI : MAXI;
    constexpr int LB = NB ? NB : MAXB;
    const T cc = (T)p.cc;
    T Cv[LI], lC[LI], Bv[LB], lB[LB];
    const T *Lyl = Lc - sy, *Lyr = Lc + sy, *Lzl = Lc + zl, *Lzr = Lc + zr;
#pragma unroll
    for (int j = 0; j < LI; ++j) {
        if (!NI && j >= p.ni) break;
        const T vc = C0[j];
        Cv[j] = vc;
        T lap = (Cp[j] + Cm[j]) * ihx2 + (Lyr[j] + Lyl[j]) * ihy2;
        if (DIM == 3) lap = lap + (Lzr[j] + Lzl[j]) * ihz2;
        lC[j] = lap - cc * vc;
    }
#pragma unroll
    for (int b = 0; b < LB; ++b) {
        if (!NB && b >= p.nb) break;
        const int g = ly_bslot<LY>(p, b);
        lB[b] = 0;
        if (ly_bface<LY>(p, b)) {
            T fv = (T)0;
#pragma unroll
            for (int f = 0; f < NF; ++f)
                if (f == g) fv = Fv[f];
            Bv[b] = face ? fv : (T)0;
        } else {
            T vc = (T)0, cxr = (T)0, cxl = (T)0;
#pragma unroll
            for (int gg = 0; gg < NG; ++gg)
                if (gg == g) {
                    vc = C0[gg];
                    cxr = Cp[gg];
                    cxl = Cm[gg];
                }
            Bv[b] = vc;
            if (ly_mobile<LY>(p, b)) {
                T lap = (cxr + cxl) * ihx2 + (Lyr[g] + Lyl[g]) * ihy2;
                if (DIM == 3) lap = lap + (Lzr[g] + Lzl[g]) * ihz2;
                lB[b] = lap - cc * vc;
            }
        }
    }
    // free buffer seen by each bound field: total minus its buffer's bound
    // fields, subtracted in bound order
    T frb[LB];
#pragma unroll
    for (int b = 0; b < LB; ++b) {
        if (!NB && b >= p.nb) break;
        T v = (T)ly_total<LY>(p, ly_bbuf<LY>(p, b), z);
#pragma unroll
        for (int b2 = 0; b2 < LB; ++b2) {
            if (!NB && b2 >= p.nb) break;
            if (ly_bbuf<LY>(p, b2) == ly_bbuf<LY>(p, b)) v -= Bv[b2];
        }
        frb[b] = v;
    }
    // R1 = koff B - kon C Bfree per bound field; the evolution of a bound
    // field is D lap - R1 (-R1 when it does not diffuse), an ion's gains the
    // sum of its fields' R1 (a z-face buffer's only at face nodes)
    T eB[LB], Rr[LI];
#pragma unroll
    for (int j = 0; j < LI; ++j) Rr[j] = (T)-0.0;  // -0 + x == x: the first term enters exactly
#pragma unroll
    for (int b = 0; b < LB; ++b) {
        if (!NB && b >= p.nb) break;
        const int j = ly_bion<LY>(p, b);
        T cj = (T)0;
#pragma unroll
        for (int jj = 0; jj < LI; ++jj)
            if (jj == j) cj = Cv[jj];
        const T R1 = (T)p.koff[b] * Bv[b] - (T)p.kon[b] * cj * frb[b];
        eB[b] = ly_mobile<LY>(p, b) ? (T)p.Db[b] * lB[b] - R1 : -R1;
        const bool on = !ly_bface<LY>(p, b) || face;
#pragma unroll
        for (int jj = 0; jj < LI; ++jj)
            if (jj == j && on) Rr[jj] += R1;
    }
#pragma unroll
    for (int j = 0; j < LI; ++j) {
        if (!NI && j >= p.ni) break;
        T e = (T)p.Dion[j] * lC[j] + (Rr[j] + J[j]);
        if ((DIM == 2 || z == p.nz - 1) && ly_sl<LY>(p, j)) {
            T s1 = (T)p.K1[j] + Cv[j];
            T s2 = (T)p.K2[j] + Cv[j];
            e *= (T)1 / ((T)1 + (T)p.N1[j] * (T)p.K1[j] / (s1 * s1) + (T)p.N2[j] * (T)p.K2[j] / (s2 * s2));
        }
        G[j] = Cv[j] + e * dt;
    }
#pragma unroll
    for (int b = 0; b < LB; ++b) {
        if (!NB && b >= p.nb) break;
        const int g = ly_bslot<LY>(p, b);
        const T v = Bv[b] + eB[b] * dt;
        if (ly_bface<LY>(p, b)) {
#pragma unroll
            for (int f = 0; f < NF; ++f)
                if (f == g) Fo[f] = v;
        } else {
#pragma unroll
            for (int gg = 0; gg < NG; ++gg)
                if (gg == g) G[gg] = v;
        }
    }
}

// Per-thread slot geometry of a plane (shared by both sweeps).
template <int DIM, int SL>
struct PlaneSlots {
    int q[SL], zl[SL], zr[SL], fz[SL], yy[SL];
    bool act[SL], inner[SL];
    __device__ __forceinline__ void init(const Params &p, int P) {
#pragma unroll
        for (int k = 0; k < SL; ++k) {
            q[k] = threadIdx.x + k * blockDim.x;
            act[k] = q[k] < P;
            const int z = DIM == 3 ? q[k] % p.nz : 0;
            yy[k] = DIM == 3 ? q[k] / p.nz : q[k];
            inner[k] = act[k] && yy[k] >= 1 && yy[k] <= p.ny - 2;
            zl[k] = DIM == 3 ? (z == 0 ? 1 : -1) : 0;
            zr[k] = DIM == 3 ? (z == p.nz - 1 ? -1 : 1) : 0;
            fz[k] = (DIM == 3 && act[k]) ? (z == 0 ? 0 : (z == p.nz - 1 ? 1 : -1)) : -1;
        }
    }
};

// Index of interior node (x, y) in the ring of nodes next to the rim (x = 1,
// x = nx-2, y = 1, y = ny-2), or -1: x = 1 first, then x = nx-2, then the
// y = 1 and y = ny-2 rows without their ends.  The cytosol box reads the
// step-k+1 and step-k values that k_sweep2 leaves there.
__host__ __device__ __forceinline__ int ring_index(int x, int y, int nx, int ny) {
    if (x == 1) return y - 1;
    if (x == nx - 2) return (ny - 2) + y - 1;
    if (y == 1) return 2 * (ny - 2) + x - 2;
    if (y == ny - 2) return 2 * (ny - 2) + (nx - 4) + x - 2;
    return -1;
}
__host__ __device__ __forceinline__ int ring_nodes(int nx, int ny) {
    return 2 * (ny - 2) + 2 * (nx > 4 ? nx - 4 : 0);
}

// First entry of a site's node-sorted channel list at or after plane x.
__device__ __forceinline__ int first_channel_at(const int *__restrict__ snode, int lo, int hi, int key) {
    while (lo < hi) {
        const int m = (lo + hi) >> 1;
        if (snode[m] < key) lo = m + 1; else hi = m;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// The state sweep, plane-marching form.  One workgroup owns one x-chunk of one
// site and walks its planes x0..x1-1 in order.  A plane (fixed x) is the
// contiguous [y][z] block of P = ny*nz nodes of every field; thread t owns the
// plane slots q = t + k*blockDim (k < SL) in every plane, so the x-neighbours
// of its nodes are its own registers (planes x-1, x, x+1 held in A, B, C; plane
// x+2 prefetched into N while x is computed).  Only the current plane goes
// through LDS (double-buffered, one LDS-only barrier per plane) for the y and
// z neighbours.  Every state value is read from HBM once per chunk; chunks add
// two halo planes.  Units are laid out so that consecutive chunks of a site
// land on the same XCD (blocks are dealt to XCDs round-robin).
// Arithmetic and operation order as k_step (DyadRDCPU.cpp:404-506).
template <typename T, int DIM, int NI, int NB, int NGT, int NFT, int LY, int SL, int PD>
__global__ __launch_bounds__(512) void k_sweep(const T *__restrict__ gin, T *__restrict__ gout,
                                               const T *__restrict__ fin, T *__restrict__ fout,
                                               const int *__restrict__ ch_off, const int *__restrict__ snode,
                                               const double *__restrict__ sJ, int nch, Params p, double dt_,
                                               int chunk, int nchunks, int units) {
    constexpr int LI = NI ? NI : MAXI;
    constexpr int LB = NB ? NB : MAXB;
    constexpr int NG = NGT ? NGT : LI + LB;         // grid fields (upper bound when NGT = 0)
    constexpr int NF = NGT ? (NFT ? NFT : 1) : LB;  // face fields (upper bound when NGT = 0)
    constexpr int RS = 3 + PD;                      // register ring: planes x-1 .. x+1+PD
    extern __shared__ unsigned char smem_raw[];
    T *lds = reinterpret_cast<T *>(smem_raw);

    const int per = gridDim.x >> 3;
    const int unit = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (unit >= units) return;
    const int s = unit / nchunks;
    const int cidx = unit - s * nchunks;
    const int x0 = 1 + cidx * chunk;
    const int x1 = min(p.nx - 1, x0 + chunk);
    const int P = p.ny * p.nz;
    const int ng = NGT ? NGT : p.n_grid, nfc = NGT ? NFT : p.n_face;  // compile-time with the layout
    const size_t fs = (size_t)p.S * p.NS;
    const T *gs = gin + (size_t)s * p.NS;
    T *os = gout + (size_t)s * p.NS;
    const T dt = (T)dt_;
    const T ihx2 = (T)p.ihx2, ihy2 = (T)p.ihy2, ihz2 = (T)p.ihz2;

    PlaneSlots<DIM, SL> ps;
    ps.init(p, P);

    // R[r] holds plane x-1+r of every grid field; F[r] the face values of plane x+r
    T R[RS][SL][NG];
    T F[PD + 1][SL][NF];
    auto load_plane = [&](T (&D)[SL][NG], int x) {
#pragma unroll
        for (int k = 0; k < SL; ++k)
#pragma unroll
            for (int g = 0; g < NG; ++g)
                if (g < ng && ps.act[k]) D[k][g] = ld_off(gs + g * fs, (unsigned)((x * P + ps.q[k]) * (int)sizeof(T)));
    };
    auto load_face = [&](T (&D)[SL][NF], int x) {
#pragma unroll
        for (int k = 0; k < SL; ++k)
#pragma unroll
            for (int f = 0; f < NF; ++f)
                if (f < nfc && ps.fz[k] >= 0)
                    D[k][f] = ld_off(fin + ((size_t)f * p.S + s) * 2 * p.nFP,
                                     (unsigned)((ps.fz[k] * p.nFP + x * p.ny + ps.yy[k]) * (int)sizeof(T)));
    };
#pragma unroll
    for (int r = 0; r < RS - 1; ++r)
        if (x0 - 1 + r <= x1) load_plane(R[r], x0 - 1 + r);
#pragma unroll
    for (int r = 0; r < PD; ++r)
        if (x0 + r < x1) load_face(F[r], x0 + r);

    int cur = first_channel_at(snode, ch_off[s], ch_off[s + 1], x0 * P);
    const int chi = ch_off[s + 1];

    for (int x = x0; x < x1; ++x) {
        if (x + 1 + PD <= x1) load_plane(R[RS - 1], x + 1 + PD);
        if (x + PD < x1) load_face(F[PD], x + PD);
        T *L = lds + (size_t)((x - x0) & 1) * ng * P;
#pragma unroll
        for (int k = 0; k < SL; ++k)
#pragma unroll
            for (int g = 0; g < NG; ++g)
                if (g < ng && ps.act[k]) L[ps.q[k] * ng + g] = R[1][k][g];
        lds_barrier();

        // channel currents on this plane (the last channel on a node wins)
        T J[SL][LI];
#pragma unroll
        for (int k = 0; k < SL; ++k)
#pragma unroll
            for (int j = 0; j < LI; ++j) J[k][j] = 0;
        const int pend = (x + 1) * P;
        while (cur < chi) {
            const int nd = snode[cur];
            if (nd >= pend) break;
            const int qq = nd - x * P;
#pragma unroll
            for (int k = 0; k < SL; ++k)
                if (ps.q[k] == qq)
#pragma unroll
                    for (int j = 0; j < LI; ++j)
                        if (NI || j < p.ni) J[k][j] = (T)sJ[(size_t)j * nch + cur];
            ++cur;
        }

        // rim nodes are written back with their values (whole cache lines per plane)
#pragma unroll
        for (int k = 0; k < SL; ++k) {
            if (!ps.act[k]) continue;
            const int c = ps.q[k];
            const int z = DIM == 3 ? c % p.nz : 0;
            T G[NG], Fo[NF];
            if (ps.inner[k]) {
                node_update<T, DIM, NI, NB, NG, NF, LY>(p, R[0][k], R[1][k], R[2][k], L + c * ng, p.nz * ng,
                                                    ps.zl[k] * ng, ps.zr[k] * ng, z,
                                                    ps.fz[k] >= 0, F[0][k], J[k], dt, ihx2, ihy2, ihz2, G, Fo);
            } else {
#pragma unroll
                for (int g = 0; g < NG; ++g) G[g] = R[1][k][g];
#pragma unroll
                for (int f = 0; f < NF; ++f) Fo[f] = F[0][k][f];
            }
            const unsigned o = (unsigned)((x * P + c) * (int)sizeof(T));
#pragma unroll
            for (int g = 0; g < NG; ++g)
                if (g < ng) st_off(os + g * fs, o, G[g]);
            if (ps.fz[k] >= 0)
#pragma unroll
                for (int f = 0; f < NF; ++f)
                    if (f < nfc)
                        st_off(fout + ((size_t)f * p.S + s) * 2 * p.nFP,
                               (unsigned)((ps.fz[k] * p.nFP + x * p.ny + ps.yy[k]) * (int)sizeof(T)), Fo[f]);
        }
        // advance the rings by one plane
#pragma unroll
        for (int r = 0; r < RS - 1; ++r)
#pragma unroll
            for (int k = 0; k < SL; ++k)
#pragma unroll
                for (int g = 0; g < NG; ++g) R[r][k][g] = R[r + 1][k][g];
#pragma unroll
        for (int r = 0; r < PD; ++r)
#pragma unroll
            for (int k = 0; k < SL; ++k)
#pragma unroll
                for (int f = 0; f < NF; ++f) F[r][k][f] = F[r + 1][k][f];
    }
}

// ---------------------------------------------------------------------------
// Two steps in one pass (temporal blocking).  The workgroup marches planes as
// k_sweep does; at plane x it computes step k+1 on plane x (sub-step A, from
// the step-k ring and the currents the channel kernel left in sJ) and step
// k+2 on plane x-1 (sub-step B, from the ring of step-k+1 planes x-2, x-1, x
// that A produced).  Only step k+2 is written: one read and one write of the
// state per two steps.  The currents of step k+2 are those the channel kernel
// would compute from the step-k+1 mouth values and the states it gated at
// step k+1 (already in state[]): J = state * g * (jSR - c_mouth), the last
// channel on a node winning; the step-k+1 values of every mouth node are
// written to snap[] for the next channel call, which derives the mouth
// snapshot (ions, bound fields, free buffers) from them and gates step k+2.  Rim nodes stay at their step-k values
// (constant Dirichlet rim: no cytosol pool); with a jSR pool the currents use
// the pool's value, which the channel call before the pass has advanced.  Chunks of one site recompute one halo plane of
// step k+1 on each side.  Node arithmetic is node_update, as in k_sweep, so
// two steps here equal two k_sweep steps bit for bit.
// The FP64 3D form with one plane slot per thread holds 81 VGPRs, one over
// what six waves per SIMD allow: the hint brings it to 80 without spills, so
// a CU holds three 7-wave workgroups (448 threads, one site each) instead of
// two (4096-site batch with two half-batch streams: 52.3 against 51.6
// simulated dyad-ms/s).  Other forms keep the compiler's choice.
#ifndef CASIM_SWEEP2_ATTR
#define CASIM_SWEEP2_ATTR __attribute__((amdgpu_waves_per_eu(SL == 1 && DIM == 3 && sizeof(T) == 8 ? 6 : 1)))
#endif
template <typename T, int DIM, int NI, int NB, int NGT, int NFT, int LY, int SL, bool RING>
__global__ __launch_bounds__(512) CASIM_SWEEP2_ATTR void k_sweep2(const T *__restrict__ gin, T *__restrict__ gout,
                                                const T *__restrict__ fin, T *__restrict__ fout,
                                                const int *__restrict__ ch_off, const int *__restrict__ snode,
                                                const double *__restrict__ sJ, const double *__restrict__ jnew,
                                                T *__restrict__ ring, const Channels a, Params p, double dt_,
                                                int chunk, int nchunks, int units, int s0) {
    constexpr int LI = NI ? NI : MAXI;
    constexpr int LB = NB ? NB : MAXB;
    constexpr int NG = NGT ? NGT : LI + LB;
    constexpr int NF = NGT ? (NFT ? NFT : 1) : LB;
    extern __shared__ unsigned char smem_raw[];
    T *lds = reinterpret_cast<T *>(smem_raw);

    // units cover the sites s0 .. s0 + units / nchunks - 1
    const int per = gridDim.x >> 3;
    const int unit = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (unit >= units) return;
    const int s = s0 + unit / nchunks;
    const int cidx = unit - (s - s0) * nchunks;
    const int x0 = 1 + cidx * chunk;
    const int x1 = min(p.nx - 1, x0 + chunk);  // step k+2 planes [x0, x1)
    const int xa0 = x0 - 1;                    // step k+1 planes [xa0, x1]
    const int P = p.ny * p.nz;
    const int ng = NGT ? NGT : p.n_grid, nfc = NGT ? NFT : p.n_face;  // compile-time with the layout
    const size_t fs = (size_t)p.S * p.NS;
    const T *gs = gin + (size_t)s * p.NS;
    T *os = gout + (size_t)s * p.NS;
    const T dt = (T)dt_;
    const T ihx2 = (T)p.ihx2, ihy2 = (T)p.ihy2, ihz2 = (T)p.ihz2;
    // LDS planes: double-buffered in 3D (one barrier per plane); single in 2D,
    // where half the LDS doubles the workgroups per CU and that pays for the
    // second barrier (measured: 2D 201 x 201 sites +10 %, 3D 81 x 81 x 5 -9 %)
    constexpr bool DB = DIM == 3;
    // with double buffers the previous plane stays in LDS for one more
    // plane; a node's own x-1 values are read from there (only the owning
    // thread writes them, so no barrier guards them) instead of a register
    // ring plane per step, which frees 12 FP64 registers per plane slot
    constexpr bool LPREV = DB;
    T *LA = lds;                                      // step-k planes
    T *LB2 = lds + (size_t)(DB ? 2 : 1) * ng * P;    // step-k+1 planes

    PlaneSlots<DIM, SL> ps;
    ps.init(p, P);
    // per-lane slot flags (bit 0: slot on the plane, bit 1: interior node,
    // bit 2: z-face node), re-read each plane through an empty asm so that
    // the compiler tests them where they are used instead of keeping one
    // 64-bit lane mask per flag live across the plane loop
    int fl[SL];
#pragma unroll
    for (int k = 0; k < SL; ++k) fl[k] = (ps.act[k] ? 1 : 0) | (ps.inner[k] ? 2 : 0) | (ps.fz[k] >= 0 ? 4 : 0);

    T R[4][SL][NG];   // step k: planes x-1, x, x+1, x+2 (prefetch)
    T F[2][SL][NF];   // step k faces: planes x, x+1 (prefetch)
    T A[3][SL][NG];   // step k+1: planes x-2, x-1, x
    T FA[2][SL][NF];  // step k+1 faces: planes x-1, x
    auto load_plane = [&](T (&D)[SL][NG], int x) {
#pragma unroll
        for (int k = 0; k < SL; ++k)
#pragma unroll
            for (int g = 0; g < NG; ++g)
                if (g < ng && (fl[k] & 1))
                    D[k][g] = ld_off(gs + g * fs, (unsigned)((x * P + ps.q[k]) * (int)sizeof(T)));
    };
    auto load_face = [&](T (&D)[SL][NF], int x) {
#pragma unroll
        for (int k = 0; k < SL; ++k)
#pragma unroll
            for (int f = 0; f < NF; ++f)
                if (f < nfc && ps.fz[k] >= 0)
                    D[k][f] = ld_off(fin + ((size_t)f * p.S + s) * 2 * p.nFP,
                                     (unsigned)((ps.fz[k] * p.nFP + x * p.ny + ps.yy[k]) * (int)sizeof(T)));
    };
    // planes xa0-1 .. xa0+1 (plane -1 does not exist: A copies the rim plane 0)
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const int xr = xa0 - 1 + r;
        if (xr >= 0 && xr <= p.nx - 1) load_plane(R[r], xr);
    }
    if (LPREV && xa0 >= 1) {
        T *Lp = LA + (size_t)((xa0 - 1) & 1) * ng * P;
#pragma unroll
        for (int k = 0; k < SL; ++k)
#pragma unroll
            for (int g = 0; g < NG; ++g)
                if (g < ng && (fl[k] & 1)) Lp[ps.q[k] * ng + g] = R[0][k][g];
    }
    load_face(F[0], xa0);

    const int chi = ch_off[s + 1];
    int curA = first_channel_at(snode, ch_off[s], chi, xa0 * P);
    int curB = first_channel_at(snode, ch_off[s], chi, x0 * P);

    // the prologue's loads land before the loop: registers entering it then
    // carry no pending load, and the compiler's wait counts inside the loop
    // cover only the plane prefetch (otherwise it waits for every load in
    // flight, the prefetch included, before the first use of a carried value)
    __builtin_amdgcn_s_waitcnt(0);
    for (int x = xa0; x <= x1; ++x) {
#pragma unroll
        for (int k = 0; k < SL; ++k) asm volatile("" : "+v"(fl[k]));
        if (x + 2 <= min(x1 + 1, p.nx - 1)) load_plane(R[3], x + 2);
        if (x + 1 <= x1) load_face(F[1], x + 1);
        const bool doB = x - 1 >= x0;
        const bool interiorA = x >= 1 && x <= p.nx - 2;
        T *La = LA + (DB ? (size_t)(x & 1) * ng * P : 0);
        T *Lb = LB2 + (DB ? (size_t)((x - 1) & 1) * ng * P : 0);
        // single buffer: every thread is done with the previous plane before it is overwritten
        if (!DB) lds_barrier();
        // with LPREV every plane goes to LDS (the rim planes and plane xa0 of
        // step k+1 are the x-1 neighbours of the next plane's update)
        const bool wA = LPREV || interiorA, wB = LPREV ? x >= x0 : doB;
#pragma unroll
        for (int k = 0; k < SL; ++k)
#pragma unroll
            for (int g = 0; g < NG; ++g)
                if (g < ng && (fl[k] & 1)) {
                    if (wA) La[ps.q[k] * ng + g] = R[1][k][g];
                    if (wB) Lb[ps.q[k] * ng + g] = A[1][k][g];
                }
        lds_barrier();

        // sub-step A: step k+1 on plane x
        {
            T J[SL][LI];
#pragma unroll
            for (int k = 0; k < SL; ++k)
#pragma unroll
                for (int j = 0; j < LI; ++j) J[k][j] = 0;
            const int pend = (x + 1) * P;
            while (curA < chi) {
                const int nd = snode[curA];
                if (nd >= pend) break;
                const int qq = nd - x * P;
#pragma unroll
                for (int k = 0; k < SL; ++k)
                    if (ps.q[k] == qq)
#pragma unroll
                        for (int j = 0; j < LI; ++j)
                            if (NI || j < p.ni) J[k][j] = (T)sJ[(size_t)j * a.nch + curA];
                ++curA;
            }
#pragma unroll
            for (int k = 0; k < SL; ++k) {
                if (interiorA && (fl[k] & 2)) {
                    const int c = ps.q[k];
                    const int z = DIM == 3 ? c % p.nz : 0;
                    const T *cm = R[0][k];
                    if constexpr (LPREV) cm = LA + (size_t)((x - 1) & 1) * ng * P + c * ng;
                    node_update<T, DIM, NI, NB, NG, NF, LY>(p, cm, R[1][k], R[2][k], La + c * ng, p.nz * ng,
                                                        ps.zl[k] * ng, ps.zr[k] * ng,
                                                        z, (fl[k] & 4) != 0, F[0][k], J[k], dt, ihx2, ihy2, ihz2,
                                                        A[2][k], FA[1][k]);
                } else {
#pragma unroll
                    for (int g = 0; g < NG; ++g) A[2][k][g] = R[1][k][g];
#pragma unroll
                    for (int f = 0; f < NF; ++f) FA[1][k][f] = F[0][k][f];
                }
            }
            // step-k+1 and step-k values of the nodes next to the rim, for a cytosol box
            if (RING && interiorA) {
                const int nzd = DIM == 3 ? p.nz : 1;
                const size_t RN = (size_t)ring_nodes(p.nx, p.ny) * nzd;
#pragma unroll
                for (int k = 0; k < SL; ++k) {
                    if (!ps.inner[k]) continue;
                    const int ri = ring_index(x, ps.yy[k], p.nx, p.ny);
                    if (ri < 0) continue;
                    const int z = DIM == 3 ? ps.q[k] % p.nz : 0;
                    T *rp = ring + (size_t)s * RN + (size_t)ri * nzd + z;
                    T *rp0 = rp + (size_t)ng * p.S * RN;  // second half: the step-k values
#pragma unroll
                    for (int g = 0; g < NG; ++g)
                        if (g < ng) {
                            rp[g * (size_t)p.S * RN] = A[2][k][g];
                            rp0[g * (size_t)p.S * RN] = R[1][k][g];
                        }
                }
            }
        }

        // The prefetched plane (and the previous plane's stores) land here,
        // before sub-step B issues this plane's stores.  Loads and stores share
        // one vector-memory counter, and a varying number of stores follows the
        // prefetch, so without this the compiler waits for everything at the
        // loop latch, this plane's stores included (1.597 -> 1.579 ms per
        // 4096-site FP64 launch).
        __builtin_amdgcn_s_waitcnt(0);
        // sub-step B: step k+2 on plane x-1
        if (doB) {
            const int xb = x - 1;
            T J[SL][LI];
#pragma unroll
            for (int k = 0; k < SL; ++k)
#pragma unroll
                for (int j = 0; j < LI; ++j) J[k][j] = 0;
            const int pend = (xb + 1) * P;
            while (curB < chi) {
                const int nd = snode[curB];
                if (nd >= pend) break;
                const int qq = nd - xb * P;
#pragma unroll
                for (int k = 0; k < SL; ++k)
                    if (ps.q[k] == qq) {
                        // the channel kernel's current, state * g * (jSR - mouth), from
                        // the step-k+1 mouth value, with state * g as the channel call
                        // before the pass left it; the node's step-k+1 values go to the
                        // snapshot the next channel call reads (CH_FROM_STORED)
#pragma unroll
                        for (int j = 0; j < LI; ++j) {
                            if (!NI && j >= p.ni) break;
                            const double ion = (double)A[1][k][j];
                            // jSR of the step: the device pool's value when pools are on
                            const double jv = jnew ? jnew[s + j * p.S] : a.pjv[(size_t)j * a.nch + curB];
                            J[k][j] = (T)(a.pg[(size_t)j * a.nch + curB] * (jv - ion));
                        }
                        double *sn = a.snap + curB;
#pragma unroll
                        for (int g = 0; g < NG; ++g)
                            if (g < ng) sn[(size_t)g * a.nch] = (double)A[1][k][g];
#pragma unroll
                        for (int f = 0; f < NF; ++f)
                            if (f < nfc) sn[(size_t)(ng + f) * a.nch] = (fl[k] & 4) ? (double)FA[0][k][f] : 0.0;
                    }
                ++curB;
            }
            // rim nodes (y = 0, ny-1) are written back with their values, so
            // that every plane's store covers whole cache lines (L2 writes
            // back partly written lines with a read-modify-write: measured
            // 1.664 against 1.594 ms per 4096-site FP64 launch)
#pragma unroll
            for (int k = 0; k < SL; ++k) {
                if (!(fl[k] & 1)) continue;
                const int c = ps.q[k];
                const int z = DIM == 3 ? c % p.nz : 0;
                T G[NG], Fo[NF];
                if (fl[k] & 2) {
                    const T *cm = A[0][k];
                    if constexpr (LPREV) cm = LB2 + (size_t)(x & 1) * ng * P + c * ng;  // plane x-2
                    node_update<T, DIM, NI, NB, NG, NF, LY>(p, cm, A[1][k], A[2][k], Lb + c * ng, p.nz * ng,
                                                        ps.zl[k] * ng, ps.zr[k] * ng, z, (fl[k] & 4) != 0,
                                                        FA[0][k], J[k], dt, ihx2, ihy2, ihz2, G, Fo);
                } else {
#pragma unroll
                    for (int g = 0; g < NG; ++g) G[g] = A[1][k][g];
#pragma unroll
                    for (int f = 0; f < NF; ++f) Fo[f] = FA[0][k][f];
                }
                const unsigned o = (unsigned)((xb * P + c) * (int)sizeof(T));
#pragma unroll
                for (int g = 0; g < NG; ++g)
                    if (g < ng) st_off(os + g * fs, o, G[g]);
                if (fl[k] & 4)
#pragma unroll
                    for (int f = 0; f < NF; ++f)
                        if (f < nfc)
                            st_off(fout + ((size_t)f * p.S + s) * 2 * p.nFP,
                                   (unsigned)((ps.fz[k] * p.nFP + xb * p.ny + ps.yy[k]) * (int)sizeof(T)), Fo[f]);
            }
        }
        // advance the rings by one plane
#pragma unroll
        for (int k = 0; k < SL; ++k) {
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                R[0][k][g] = R[1][k][g];
                R[1][k][g] = R[2][k][g];
                R[2][k][g] = R[3][k][g];
                A[0][k][g] = A[1][k][g];
                A[1][k][g] = A[2][k][g];
            }
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                F[0][k][f] = F[1][k][f];
                FA[0][k][f] = FA[1][k][f];
            }
        }
    }
}



// ---------------------------------------------------------------------------
// Kinetic rates (SheepCannell.cpp:12-18, TwoStateModel.cpp:4-8,
// TwoStatePoModel.cpp:3-8, TwoStateMWC.cpp:15-23, TwoStateaHTG.cpp:18-29).
__device__ __forceinline__ double channel_rate(int model, const double *q, int state, double ca, double mg) {
    if (model == 1) {
        if (state) return q[3] * pow(ca * 1e-3, q[4]) * 1e-3;
        double v = q[0] * pow(ca * 1e-3, q[1]);
        return ((q[2] < v) ? q[2] : v) * 1e-3;
    }
    if (model == 2) {
        double fCa = q[0], KCa = q[1], fMg = q[2], KMg = q[3], KO0 = q[4], KMgI = q[5], kclose = q[6];
        if (state) return kclose;
        double inac = KMgI * KMgI / (KMgI * KMgI + mg * mg);
        double open = pow(ca + KCa * fCa * (1 + mg / (fMg * KMg)), 4.0);
        double close = KO0 * pow(fCa * (ca + KCa * (1 + mg / KMg)), 4.0);
        double P = inac * open / (open + close);
        return kclose * P / (1 - P);
    }
    if (model == 3) {
        double fCa = q[0], KCa = q[1], fMg = q[2], KMg = q[3], KO0 = q[4], KMgI = q[5];
        double KCL = q[6], KI = q[7], KOL = q[8], kclose = q[9];
        if (state) return kclose;
        double mult = KMgI * KMgI / (KMgI * KMgI + mg * mg);
        double up = (KCL * KI * (pow(ca, 4.0) * pow(fMg, 4.0) * pow(fCa, 4.0) * pow(KMg, 4.0) * KO0 +
                                 pow(ca * fMg * KMg + fCa * (mg + fMg * KMg) * KCa, 4.0) * KOL));
        double down =
            (4 * pow(ca, 3.0) * pow(fMg, 3.0) * fCa * pow(KMg, 3.0) * KCa * KCL * KI *
                 (mg + fMg * KMg + fMg * pow(fCa, 3.0) * (mg + KMg) * KO0) * KOL +
             6 * pow(ca, 2.0) * pow(fMg, 2.0) * pow(fCa, 2.0) * pow(KMg, 2.0) * pow(KCa, 2.0) * KCL * KI *
                 (pow(mg + fMg * KMg, 2.0) + pow(fMg, 2.0) * pow(fCa, 2.0) * pow(mg + KMg, 2.0) * KO0) * KOL +
             4 * ca * fMg * pow(fCa, 3.0) * KMg * pow(KCa, 3.0) * KCL * KI *
                 (pow(mg + fMg * KMg, 3.0) + pow(fMg, 3.0) * fCa * pow(mg + KMg, 3.0) * KO0) * KOL +
             pow(fCa, 4.0) * pow(KCa, 4.0) * KCL * KI *
                 (pow(mg + fMg * KMg, 4.0) + pow(fMg, 4.0) * pow(mg + KMg, 4.0) * KO0) * KOL +
             pow(ca, 4.0) * pow(fMg, 4.0) * pow(KMg, 4.0) *
                 (KCL * KI * KOL + pow(fCa, 4.0) * KO0 * (1 + KI + KCL * KI + KCL * KI * KOL)));
        double P = mult * up / down;
        return kclose * P / (1 - P);
    }
    return 0.0;
}

// The rate with the model fixed at compile time (MODEL > 0) or read per
// channel type (MODEL = 0).
template <int MODEL>
__device__ __forceinline__ double channel_rate_m(int model, const double *q, int state, double ca, double mg) {
    return channel_rate(MODEL ? MODEL : model, q, state, ca, mg);
}

// Draw number p (0-based) after state S: jump 2p steps, then one U01d draw.
__device__ __forceinline__ double draw_at(const uint32_t *__restrict__ jump, const uint32_t S[6], int p) {
    const uint32_t *J = jump + (size_t)p * 18;
    uint32_t A1[9], A2[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        A1[i] = J[i];
        A2[i] = J[9 + i];
    }
    uint32_t st[6];
    mrg::matvec(A1, S, st, mrg::C1, mrg::M1);
    mrg::matvec(A2, S + 3, st + 3, mrg::C2, mrg::M2);
    return mrg::draw(st);
}

// ---------------------------------------------------------------------------
// Channel kernel, one 64-lane wave per site, lanes over channels.
//  CH_SNAPSHOT  mouth snapshot of ions, bound and free buffers
//               (DyadRDCPU.cpp:371-390) from the current grid
//  CH_CURRENTS  J = state * g (jSR - mouth) per ion (PassiveIonChannel.cpp:28-34,
//               DyadRDCPU.cpp:250-271) into the node-sorted current list, and
//               the site total (ion 0 restarts from zero, other ions keep
//               accumulating, as DyadRDCPU.cpp:254,260 do) when CH_ACCUMULATE
//  CH_GATE      DyadChannels::RunMC (DyadChannels.cpp:97-130) on the fresh
//               snapshot; CH_GATE_STORED on the stored one.
// The stream is consumed in channel order: one draw per channel, one more
// when it toggles.  Lanes evaluate their draw speculatively at position
// (draws so far + lanes before them) by jumping the step-start state with
// A^(2p); a ballot finds the first toggling channel, positions after it are
// re-evaluated with the extra draw accounted for, until a ballot is empty.
template <typename T, int MODEL, int NI_, int NB_, int NU_>
__global__ __launch_bounds__(64) void k_channels(const T *__restrict__ g, const T *__restrict__ f, Channels a,
                                                 Params p, Pools q, int mode, double dt, int s0) {
    // compile-time species counts where the layout is known (0: up to the maxima, guarded)
    constexpr int LI = NI_ ? NI_ : MAXI, LB = NB_ ? NB_ : MAXB, LU = NU_ ? NU_ : MAXU;
    __shared__ double Jl[LI][64];
    const int s = s0 + blockIdx.x;  // one workgroup per site of the range s0 ..
    const int lane = threadIdx.x;
    const int off = a.off[s];
    const int n = a.off[s + 1] - off;
    const size_t fs = (size_t)p.S * p.NS;
    uint32_t S6[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) S6[i] = a.rng[s * 6 + i];
    int pos = 0;
    int nopen = 0;
    double acc[LI];
#pragma unroll
    for (int k = 0; k < LI; ++k)
        acc[k] = (k == 0 || !(mode & CH_ACCUMULATE)) ? 0.0 : (k < p.ni ? a.totals[s + k * p.S] : 0.0);

    for (int c0 = 0; c0 < n; c0 += 64) {
        const int i = c0 + lane;
        const bool valid = i < n;
        int st = valid ? a.state[off + i] : 0;
        const int typ = valid ? a.type[off + i] : 0;
        double ion[LI], bnd[LB], fre[LU];
        if (mode & (CH_SNAPSHOT | CH_CURRENTS | CH_GATE | CH_FROM_STORED)) {
            const int node = valid ? a.node[off + i] : 0;
            const int z = p.dim == 3 ? node % p.nz : 0;
            const int xy = p.dim == 3 ? node / p.nz : node;
            const size_t c = (size_t)s * p.NS + node;
            // CH_FROM_STORED: the node's values that k_sweep2 stored (node-sorted position)
            const double *sn = a.snap + off + (valid ? a.perm[off + i] : 0);
            const bool stored = (mode & CH_FROM_STORED) != 0;
#pragma unroll
            for (int j = 0; j < LI; ++j)
                ion[j] = (j < p.ni) ? (stored ? sn[(size_t)j * a.nch] : (double)g[j * fs + c]) : 0.0;
#pragma unroll
            for (int b = 0; b < LB; ++b) {
                if (b >= p.nb) {
                    bnd[b] = 0.0;
                } else if (p.bface[b]) {
                    bool fc = p.dim == 3 && (z == 0 || z == p.nz - 1);
                    bnd[b] = fc ? (stored ? sn[(size_t)(p.n_grid + p.bslot[b]) * a.nch]
                                          : (double)load_face(f, p, p.bslot[b], s, z == 0 ? 0 : 1, xy))
                                : 0.0;
                } else {
                    bnd[b] = stored ? sn[(size_t)p.bslot[b] * a.nch] : (double)g[p.bslot[b] * fs + c];
                }
            }
#pragma unroll
            for (int u = 0; u < LU; ++u) fre[u] = u < p.nu ? total_at(p, u, z) : 0.0;
#pragma unroll
            for (int b = 0; b < LB; ++b)
                if (b < p.nb)
#pragma unroll
                    for (int u = 0; u < LU; ++u)
                        if (u == p.bbuf[b]) fre[u] -= bnd[b];
            if ((mode & (CH_SNAPSHOT | CH_FROM_STORED)) && valid) {
#pragma unroll
                for (int k = 0; k < LI; ++k)
                    if (k < p.ni) a.mouths[i + (size_t)a.M * (s + (size_t)k * p.S)] = ion[k];
#pragma unroll
                for (int k = 0; k < LB; ++k)
                    if (k < p.nb) a.mouths[i + (size_t)a.M * (s + (size_t)(k + p.ni) * p.S)] = bnd[k];
#pragma unroll
                for (int k = 0; k < LU; ++k)
                    if (k < p.nu) a.mouths[i + (size_t)a.M * (s + (size_t)(k + p.ni + p.nb) * p.S)] = fre[k];
            }
        }
        if (mode & CH_CURRENTS) {
            const int sp = valid ? a.perm[off + i] : 0;
#pragma unroll
            for (int k = 0; k < LI; ++k) {
                double J = 0.0;
                if (k < p.ni && valid) {
                    const double jv = q.on ? q.jnew[s + k * p.S] : a.jsr[i + (size_t)a.M * (s + (size_t)k * p.S)];
                    double cond = a.g[typ * MAXI + k] * (jv - ion[k]);
                    J = (double)st * cond;
                    a.sJ[(size_t)k * a.nch + off + sp] = J;
                    if (!q.on) a.pjv[(size_t)k * a.nch + off + sp] = jv;
                }
                Jl[k][lane] = J;
            }
            __syncthreads();
            if (mode & CH_ACCUMULATE) {
                const int m = (n - c0) < 64 ? (n - c0) : 64;
#pragma unroll
                for (int k = 0; k < LI; ++k)
                    if (k < p.ni)
                        for (int q = 0; q < m; ++q) acc[k] += Jl[k][q];
            }
            __syncthreads();
        }
        if (mode & (CH_GATE | CH_GATE_STORED)) {
            double ca, mg;
            if (mode & CH_GATE_STORED) {
                ca = (a.ca >= 0 && valid) ? a.mouths[i + (size_t)a.M * (s + (size_t)a.ca * p.S)] : 0.0;
                mg = (a.mg >= 0 && valid) ? a.mouths[i + (size_t)a.M * (s + (size_t)a.mg * p.S)] : 0.0;
            } else {
                ca = 0.0;
                mg = 0.0;
#pragma unroll
                for (int j = 0; j < LI; ++j) {
                    if (j == a.ca) ca = ion[j];
                    if (j == a.mg) mg = ion[j];
                }
            }
            const double thr = valid ? channel_rate_m<MODEL>(a.model[typ], a.par + typ * NPAR, st, ca, mg) * dt : -1.0;
            const int end = (n - c0) < 64 ? n : c0 + 64;
            int first = c0;
            while (first < end) {
                const bool active = valid && i >= first;
                double u = active ? draw_at(a.jump, S6, pos + (i - first)) : 2.0;
                const bool tr = active && (u < thr);
                const unsigned long long m = __ballot(tr);
                if (m == 0ull) {
                    pos += end - first;
                    break;
                }
                const int jl = __builtin_ctzll(m);
                if (lane == jl) st ^= 1;
                pos += (c0 + jl - first) + 2;
                first = c0 + jl + 1;
            }
            if (valid) {
                a.state[off + i] = st;
                const int sp = a.perm[off + i];
#pragma unroll
                for (int k = 0; k < LI; ++k)
                    if (k < p.ni) a.pg[(size_t)k * a.nch + off + sp] = (double)st * a.g[typ * MAXI + k];
            }
            nopen += __popcll(__ballot(valid && st == 1));
        }
    }
    if ((mode & CH_CURRENTS) && (mode & CH_ACCUMULATE) && lane == 0) {
        for (int k = 0; k < p.ni; ++k) a.totals[s + k * p.S] = acc[k];
    }
    if (q.on && lane == 0) {
        if (mode & CH_POOL_UPDATE) pool_update_site(q, s, p.S, p.ni, a.totals);
        if (mode & CH_POOL_ADVANCE) pool_advance_site(q, s, p.S, p.ni, p.nb, dt);
    }
    if ((mode & (CH_GATE | CH_GATE_STORED)) && lane == 0) {
        double *st = a.stats + (size_t)s * NSTAT;
        const double steps = st[ST_STEPS] + 1.0;
        const double t = steps * dt;
        const double nn = (double)nopen;
        if (nn != st[ST_NOPEN]) st[ST_TTE] = t;
        if (nn > st[ST_AMP]) {
            st[ST_AMP] = nn;
            st[ST_TTP] = t;
        }
        st[ST_Q] += nn * dt;
        st[ST_NOPEN] = nn;
        st[ST_STEPS] = steps;
        // time histogram input (IndChannelCREAnalyzer.cpp:69-84 on per-step records):
        // bin k >= 1 holds N after step k*every - 1, bin 0 the state after InitOpening
        if (a.trace) {
            const long long m1 = (long long)steps + 1;
            if (m1 % a.trace_every == 0 && m1 / a.trace_every < a.trace_bins)
                a.trace[(size_t)(m1 / a.trace_every) * p.S + s] = (uint8_t)(nopen < 254 ? nopen : 254);
        }
        const uint32_t *J = a.jump + (size_t)pos * 18;
        uint32_t out[6];
        mrg::matvec(J, S6, out, mrg::C1, mrg::M1);
        mrg::matvec(J + 9, S6 + 3, out + 3, mrg::C2, mrg::M2);
#pragma unroll
        for (int q = 0; q < 6; ++q) a.rng[s * 6 + q] = out[q];
    }
}

// DyadChannels::InitOpening (DyadChannels.cpp:83-95): n+1 draws, all closed,
// then `channel` open.  One thread per site.
static __global__ void k_init_opening(Channels a, int S, int site, int channel, const int *__restrict__ chans) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S || (site >= 0 && s != site)) return;
    const int off = a.off[s], n = a.off[s + 1] - off;
    const int ch = chans ? chans[s] : channel;
    for (int k = 0; k < n; ++k) a.state[off + k] = (k == ch) ? 1 : 0;
    double *st = a.stats + (size_t)s * NSTAT;
    const double n0 = (ch >= 0 && ch < n) ? 1.0 : 0.0;
    for (int i = 0; i < NSTAT; ++i) st[i] = 0.0;
    st[ST_NOPEN] = n0;
    st[ST_AMP] = n0;
    if (a.trace) a.trace[s] = (uint8_t)n0;
    uint32_t S6[6], out[6];
    for (int i = 0; i < 6; ++i) S6[i] = a.rng[s * 6 + i];
    const uint32_t *J = a.jump + (size_t)(n + 1) * 18;
    mrg::matvec(J, S6, out, mrg::C1, mrg::M1);
    mrg::matvec(J + 9, S6 + 3, out + 3, mrg::C2, mrg::M2);
    for (int i = 0; i < 6; ++i) a.rng[s * 6 + i] = out[i];
}

// SRPool::Reset (SRPool.cpp:103-116) and CytosolPool::Reset
// (CytosolPool.cpp:290-315); the jSR's NSR copy starts at the jSR value
// (SRPool.cpp:67-72).  One thread per site.
static __global__ void k_pools_reset(Pools q, int S, int ni, const double *__restrict__ jc0) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    for (int i = 0; i < ni; ++i) {
        q.jions[s + i * S] = q.jnew[s + i * S] = jc0[i];
        q.IjSR[s + i * S] = 0.0;
    }
    if (!q.cyto) return;
    int kd = 0;
    for (int j = 0; j < ni; ++j) {
        q.cions[s + j * S] = q.cbg[j];
        for (int u = 0; u < q.nbc; ++u) {
            if (j == 0) q.cfree[s + u * S] = q.ctot[u];
            const int b = q.cbind[j * q.nbc + u];
            if (b >= 0) {
                q.cbuf[s + b * S] = q.cinit[b];
                if (q.cb_dyad[b]) q.cbb[s + (kd++) * S] = q.cinit[b];
                q.cfree[s + u * S] -= q.cinit[b];
            }
        }
    }
}

// Reset (DyadRDCPU.cpp:203-229) of both ping-pong copies.
template <typename T>
__global__ void k_reset(T *g0, T *g1, T *f0, T *f1, Params p) {
    const size_t fs = (size_t)p.S * p.NS;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < fs) {
        for (int j = 0; j < p.ni; ++j) g0[j * fs + idx] = g1[j * fs + idx] = (T)p.c0[j];
        for (int b = 0; b < p.nb; ++b)
            if (!p.bface[b]) g0[p.bslot[b] * fs + idx] = g1[p.bslot[b] * fs + idx] = (T)p.binit[b];
    }
    const size_t ff = (size_t)p.S * 2 * p.nFP;
    if (idx < ff)
        for (int b = 0; b < p.nb; ++b)
            if (p.bface[b]) f0[p.bslot[b] * ff + idx] = f1[p.bslot[b] * ff + idx] = (T)p.binit[b];
}

// _SetBoundaries (DyadRDCPU.cpp:290-320) on both copies.
template <typename T>
__global__ void k_rim(T *g0, T *g1, T *f0, T *f1, const double *__restrict__ rim_ions,
                      const double *__restrict__ rim_bound, Params p) {
    const int s = blockIdx.y;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= p.NS) return;
    const int z = p.dim == 3 ? n % p.nz : 0;
    const int t = p.dim == 3 ? n / p.nz : n;
    const int y = t % p.ny, x = t / p.ny;
    if (!(x == 0 || x == p.nx - 1 || y == 0 || y == p.ny - 1)) return;
    const size_t fs = (size_t)p.S * p.NS;
    const size_t c = (size_t)s * p.NS + n;
    for (int j = 0; j < p.ni; ++j) g0[j * fs + c] = g1[j * fs + c] = (T)rim_ions[s + j * p.S];
    for (int b = 0; b < p.nb; ++b) {
        const T v = (T)rim_bound[s + b * p.S];
        if (!p.bface[b]) {
            g0[p.bslot[b] * fs + c] = g1[p.bslot[b] * fs + c] = v;
        } else if (p.dim == 3 && (z == 0 || z == p.nz - 1)) {
            const size_t q = ((size_t)p.bslot[b] * p.S + s) * 2 * p.nFP + (size_t)(z == 0 ? 0 : 1) * p.nFP + t;
            f0[q] = f1[q] = v;
        }
    }
}

// _GetGradients (DyadRDCPU.cpp:322-369, DyadRD2DCPU.cpp:305-350): the same
// one-sided rim terms, one wave per (site, field).  The sum is regrouped (lane
// partial sums, then a shuffle tree), so it matches the reference's serial sum
// to rounding, and is identical from call to call.
template <typename T>
__global__ __launch_bounds__(64) void k_efflux(const T *__restrict__ g, double *__restrict__ out, Params p, double gm,
                                               double sxm, double sym) {
#pragma clang fp contract(off)
    // one wave per (site, field): lanes take rim terms in turn, then a fixed
    // shuffle tree adds the lane sums (same order on every call)
    const int nf = p.ni + p.nb;
    const int s = blockIdx.x / nf, fi = blockIdx.x - (blockIdx.x / nf) * nf;
    const int lane = threadIdx.x;
    double D;
    const T *u;
    const size_t fs = (size_t)p.S * p.NS;
    if (fi < p.ni) {
        D = p.Dion[fi];
        u = g + fi * fs + (size_t)s * p.NS;
    } else {
        const int b = fi - p.ni;
        D = p.Db[b];
        if (p.bface[b]) {
            if (lane == 0) out[fi + nf * s] = 0.0;  // membrane buffers do not diffuse
            return;
        }
        u = g + p.bslot[b] * fs + (size_t)s * p.NS;
    }
    const int zmax = p.dim == 3 ? p.nz - 1 : 1;
    const int nz = p.nz, ny = p.ny, nx = p.nx;
    const double tx = D * gm * sxm, ty = D * gm * sym;
    double acc = 0.0;
    // x-rims: x = 0 and nx-1, y in [1, ny-2], z in [0, zmax)
    const int per_x = (ny - 2) * zmax;
    for (int t = lane; t < 2 * per_x; t += 64) {
        const int side = t >= per_x;
        const int r = t - side * per_x;
        const int y = 1 + r / zmax, z = r - (r / zmax) * zmax;
        const int x = side ? nx - 1 : 0, d = side ? -1 : 1;
        const double u1 = (double)u[z + nz * (y + ny * (x + d))];
        const double u0 = (double)u[z + nz * (y + ny * x)];
        const double u2 = (double)u[z + nz * (y + ny * (x + 2 * d))];
        acc += tx * (2 * u1 - 1.5 * u0 - 0.5 * u2);
    }
    // y-rims: y = 0 and ny-1, x in [0, nx), z in [0, zmax)
    const int per_y = nx * zmax;
    for (int t = lane; t < 2 * per_y; t += 64) {
        const int side = t >= per_y;
        const int r = t - side * per_y;
        const int x = r / zmax, z = r - (r / zmax) * zmax;
        const int y = side ? ny - 1 : 0, d = side ? -1 : 1;
        const double u1 = (double)u[z + nz * (y + d + ny * x)];
        const double u0 = (double)u[z + nz * (y + ny * x)];
        const double u2 = (double)u[z + nz * (y + 2 * d + ny * x)];
        acc += ty * (2 * u1 - 1.5 * u0 - 0.5 * u2);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if (lane == 0) out[fi + nf * s] = acc;
}

}  // namespace casim
