// qs.hip -- quasi-stationary release sites on the device, with the mouth
// [Ca2+] from the surrogate ensemble (include/casim_surrogate.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/casim_surrogate.h"
#include "dyad_internal.hpp"

using namespace casim;

namespace casim {

constexpr int QS_NEAR = 16;
constexpr int QS_F = 8 + 3 * QS_NEAR;  // geometric features per RyR
constexpr int QS_FMAX = QS_F + 1;      // + the linear-superposition feature
constexpr int QS_MAXCH = 128;
constexpr int QS_MAXMODES = 8;  // stationary model: eigenmodes (buffers + 1)

// Per-site statistics: the event record's columns (QS_AMP .. QS_EVENTS) and
// the same quantities cut at t_max (QS_W*), for comparison with fixed-step runs.
enum QsStat {
    QS_AMP = 0, QS_TTP, QS_TTE, QS_Q, QS_TCUR, QS_NCUR, QS_EVENTS,
    QS_WAMP, QS_WTTE, QS_WQ, QS_WN,
    QS_NSTAT
};

struct QsDev {
    int S;
    const int *off;
    const float *x, *y;  // lattice units relative to the cleft centre
    float half;          // half extent in lattice units
    int *state;
    uint32_t *rng;
    double *t;
    int *active;   // 1 while the site evolves
    int *waiting;  // 0 no, 1 needs the full solver, 2 values supplied
    int *crow;     // [S] first surrogate row of the site this round, -1 if none
    int *list;     // the sites with rows this round, in site order (k_qs_plan)
    uint8_t *trace;  // [bins][S] open channels per time bin (or null)
    int *cur_bin;    // [S] first bin not yet written
    double bin_ms;
    int bins;
    int R;         // surrogate rows this round (member stride of out)
    double *stats;
    double *jsr;      // [S] current jSR
    double *jsr_min;  // [S] lowest jSR over [0, t_max]
    int deplete;
    double Vr, cq, ncq, kcq, tau, nsr, gch;
    float *X;        // [channels x nf]
    int nf;          // features per row: QS_F, or QS_F + 1 with the superposition table
    const float *lin_tab;  // [lin_n] added [Ca2+] (uM) against r^2 (lattice units^2), at jSR lin_jsr
    int lin_n;
    int lin_residual;  // the ensemble predicts ln [Ca2+] minus feature QS_F
    float lin_bg, lin_jsr;
    float *out;      // [members x channels]
    float *ca;       // [channels] mouth [Ca2+] in use
    float *spread;   // [channels]
    float *ovr;      // [channels] supplied by the full solver
    // the reference's stationary model (LinearNaraghiNeher + StationaryDiffusionModel)
    // in place of the ensemble when st_n > 0
    const double *xd, *yd;  // [channels] positions in the coordinate file's units
    int st_n;
    double st_coef[QS_MAXMODES], st_decay[QS_MAXMODES], st_bg, st_open;
    int model;
    double par[NPAR];
    double mg;
    double t_max;
    float thresh;
    int members;
    float ym, ys;
};

// Added [Ca2+] of one open RyR at squared distance r2 (lattice units): the
// table, linearly interpolated in r^2, decaying as 1/r past its end.
__device__ __forceinline__ float lin_added(const QsDev &q, float r2) {
    const float top = (float)(q.lin_n - 1);
    if (r2 >= top) return q.lin_tab[q.lin_n - 1] * sqrtf(top / r2);
    const int k = (int)r2;
    const float w = r2 - (float)k;
    return q.lin_tab[k] + w * (q.lin_tab[k + 1] - q.lin_tab[k]);
}

// StationaryDiffusionModel::GetConcentrations (StationaryDiffusionModel.cpp:3-11)
// for channel i of a site: over the open channels j in order, the [Ca2+]
// LinearNaraghiNeher's table holds for their distance (LinearNaraghiNeher.cpp:
// 62-71: the modes' sum over r, "Open Concentration" at r <= 1e-6), then the
// background.  FP64 throughout, as the reference.
__device__ double stationary_mouth(const QsDev &q, int off, int n, int i) {
#pragma clang fp contract(off)
    const double xi = q.xd[off + i], yi = q.yd[off + i];
    double c = 0.0;
    for (int j = 0; j < n; ++j) {
        if (q.state[off + j] != 1) continue;  // isOpen() * added adds +0
        const double dx = q.xd[off + j] - xi, dy = q.yd[off + j] - yi;
        const double r = sqrt(dx * dx + dy * dy + 0.0);  // Point::d, z = 0
        double a = 0.0;
        if (r > 1e-6) {
            for (int m = 0; m < q.st_n; ++m) a += 1.0 / r * q.st_coef[m] * exp(-r * q.st_decay[m]);
        } else {
            a = q.st_open;
        }
        c += a;
    }
    return c + q.st_bg;
}

// Nearest-open list of one channel: the QS_NEAR nearest open channels as
// keys (bits of r << 32 | channel index), ascending.  r >= 0, whose bits
// order like its values, so the key order is r's with ties broken by the
// channel index: the order of a stable argsort over the channels in index
// order.  Kept by insertion; one compare and two selects per step.
struct NearList {
    unsigned long long key[QS_NEAR];
    __device__ __forceinline__ void clear() {
#pragma unroll
        for (int k = 0; k < QS_NEAR; ++k) key[k] = ~0ull;
    }
    __device__ __forceinline__ bool insert(unsigned long long kk) {
        if (!(kk < key[QS_NEAR - 1])) return false;
        key[QS_NEAR - 1] = kk;
#pragma unroll
        for (int k = QS_NEAR - 1; k > 0; --k) {
            const unsigned long long lo = key[k] < key[k - 1] ? key[k] : key[k - 1];
            const unsigned long long hi = key[k] < key[k - 1] ? key[k - 1] : key[k];
            key[k - 1] = lo;
            key[k] = hi;
        }
        return true;
    }
    static __device__ __forceinline__ unsigned long long make(float r, int j) {
        return ((unsigned long long)__float_as_uint(r) << 32) | (unsigned)j;
    }
};

// Features of every channel of every active site, one workgroup of four
// waves per site: lanes over channels (64-channel chunks), and when many
// channels are open the open list is split into consecutive parts, one wave
// each.  The wave of part 0 then adds the other parts' sums in part order
// and inserts their nearest lists in part order, which gives the stable
// nearest list of the whole scan (the sums regroup, to rounding).
constexpr int QS_FWAVES = 4;
constexpr int QS_PSTRIDE = 4 + 2 * QS_NEAR;  // sinv, sexp, lin, a pad word, then the keys (8-byte aligned)
__global__ __launch_bounds__(64 * QS_FWAVES) void k_qs_features(QsDev q) {
    __shared__ float sx[QS_MAXCH], sy[QS_MAXCH];
    __shared__ int so[QS_MAXCH], ol[QS_MAXCH];
    __shared__ int s_nopen;
    __shared__ __attribute__((aligned(16))) float part[QS_FWAVES - 1][64 * QS_PSTRIDE];
    const int s = q.list[blockIdx.x];  // one workgroup per site with rows this round
    const int base = q.crow[s];
    if (base < 0) return;
    const int off = q.off[s], n = q.off[s + 1] - off;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    // the open channels in index order (a ballot per 64 channels)
    if (w == 0) {
        int nopen = 0;
        for (int c0 = 0; c0 < n; c0 += 64) {
            const int i = c0 + lane;
            const bool op = i < n && q.state[off + i] == 1;
            if (i < n) {
                sx[i] = q.x[off + i];
                sy[i] = q.y[off + i];
                so[i] = op;
            }
            const unsigned long long m = __ballot(op);
            if (op) ol[nopen + __popcll(m & ((1ull << lane) - 1ull))] = i;
            nopen += __popcll(m);
        }
        if (lane == 0) s_nopen = nopen;
    }
    __syncthreads();
    const int nopen = s_nopen;
    const int nchunk = (n + 63) / 64;  // 1 or 2
    const int parts = nopen >= 32 ? QS_FWAVES / nchunk : 1;
    const int c = w / parts, h = w - c * parts;
    const int i = c * 64 + lane;
    const bool mine = c < nchunk && i < n;
    const int len = (nopen + parts - 1) / parts;
    const int o0 = h * len, o1 = min(nopen, o0 + len);
    NearList nl;
    nl.clear();
    float sinv = 0.f, sexp = 0.f, lin = 0.f;
    const float xi = mine ? sx[i] : 0.f, yi = mine ? sy[i] : 0.f;
    if (mine) {
        for (int o = o0; o < o1; ++o) {
            const int j = ol[o];
            if (j == i) continue;
            const float dx = sx[j] - xi, dy = sy[j] - yi;
            const float r2 = dx * dx + dy * dy;
            const float r = sqrtf(r2);
            sinv += 1.f / r;
            sexp += expf(-r);
            if (q.lin_n) lin += lin_added(q, r2);
            nl.insert(NearList::make(r, j));
        }
    }
    if (parts > 1) {
        if (h > 0 && mine) {
            float *pp = part[c * (parts - 1) + h - 1] + lane * QS_PSTRIDE;
            pp[0] = sinv;
            pp[1] = sexp;
            pp[2] = lin;
            unsigned long long *pk = reinterpret_cast<unsigned long long *>(pp + 4);
#pragma unroll
            for (int k = 0; k < QS_NEAR; ++k) pk[k] = nl.key[k];
        }
        __syncthreads();
        if (h == 0 && mine) {
            for (int hh = 1; hh < parts; ++hh) {
                const float *pp = part[c * (parts - 1) + hh - 1] + lane * QS_PSTRIDE;
                sinv += pp[0];
                sexp += pp[1];
                lin += pp[2];
                // the part's list is ascending: the first entry that does not
                // enter ends it
                const unsigned long long *pk = reinterpret_cast<const unsigned long long *>(pp + 4);
                for (int k = 0; k < QS_NEAR; ++k)
                    if (!nl.insert(pk[k])) break;
            }
        }
    }
    if (h != 0 || !mine) return;
    float *f = q.X + (size_t)(base + i) * q.nf;
    f[0] = (float)so[i];
    f[1] = (float)(q.jsr[s] / 1000.0);
    f[2] = nopen / 100.f;
    f[3] = sinv;
    f[4] = sexp;
    f[5] = xi;
    f[6] = yi;
    f[7] = q.half - fmaxf(fabsf(xi), fabsf(yi));
#pragma unroll
    for (int k = 0; k < QS_NEAR; ++k) {
        const bool ok = nl.key[k] != ~0ull;
        const int j = ok ? (int)(unsigned)nl.key[k] : i;
        const float r = __uint_as_float((unsigned)(nl.key[k] >> 32));
        f[8 + 3 * k] = ok ? sx[j] - xi : 0.f;
        f[9 + 3 * k] = ok ? sy[j] - yi : 0.f;
        f[10 + 3 * k] = ok ? 1.f / r : 0.f;
    }
    if (q.lin_n) {
        // ln of the superposition estimate: background + jSR-scaled sum over open RyRs (self included)
        if (so[i]) lin += q.lin_tab[0];
        f[QS_F] = logf(q.lin_bg + (float)(q.jsr[s] / (double)q.lin_jsr) * lin);
    }
}

// SRPool::Run (SRPool.cpp:70-89) over an interval dt with a fixed open set:
// I(c) = c * geff (mouth [Ca2+] proportional to c within the interval),
// explicit steps of at most 1 us, operations in the pool's order.
__device__ double jsr_advance(const QsDev &q, double c, double geff, double dt) {
#pragma clang fp contract(off)
    double n = ceil(dt / 1e-3);
    if (n < 1.0) n = 1.0;
    if (n > 1e6) n = 1e6;
    const double h = dt / n;
    for (double k = 0; k < n; k += 1.0) {
        const double kc = q.kcq + c;
        const double beta = 1 + q.ncq * q.kcq * q.cq / (kc * kc);
        const double Iref = (q.nsr - c) / q.tau;
        c = c + h / beta * (Iref - c * geff * q.Vr);
    }
    return c;
}

// Ensemble mean / spread of every channel of active sites, then one Gillespie
// event per site unless the surrogate is too uncertain for it: the site asks
// for the full solver when the rate-weighted mean of the relative spreads,
// sum_i (rate_i / total) spread_i, exceeds the threshold.
__global__ __launch_bounds__(64) void k_qs_event(QsDev q) {
    __shared__ double rate[QS_MAXCH], wv[QS_MAXCH], pcv[QS_MAXCH];
    __shared__ float spv[QS_MAXCH], cav[QS_MAXCH];
    __shared__ double s_total;
    const int s = q.list[blockIdx.x];  // one workgroup per site with rows this round
    const int lane = threadIdx.x;
    const int base = q.crow[s];
    if (base < 0) return;
    const int off = q.off[s], n = q.off[s + 1] - off;
    const int wait = q.waiting[s];
    for (int i = lane; i < n; i += 64) {
        const size_t row = off + i;
        float ca, sp;
        if (q.st_n > 0) {
            // the reference's stationary model: rates from the FP64 value
            const double cd = stationary_mouth(q, off, n, i);
            q.ca[row] = (float)cd;
            q.spread[row] = 0.f;
            spv[i] = 0.f;
            cav[i] = (float)cd;
            rate[i] = channel_rate(q.model, q.par, q.state[row], cd, q.mg);
            continue;
        }
        if (wait == 2) {
            ca = q.ovr[row];
            sp = 0.f;
        } else {
            float sum = 0.f, sum2 = 0.f;
            const float res = q.lin_residual ? q.X[(size_t)(base + i) * q.nf + QS_F] : 0.f;
            for (int e = 0; e < q.members; ++e) {
                const float v = expf(q.out[(size_t)e * q.R + base + i] * q.ys + q.ym + res);
                sum += v;
                sum2 += v * v;
            }
            ca = sum / q.members;
            const float var = fmaxf(0.f, sum2 / q.members - ca * ca);
            sp = ca > 0.f ? sqrtf(var) / ca : 0.f;
        }
        q.ca[row] = ca;
        q.spread[row] = sp;
        spv[i] = sp;
        cav[i] = ca;
        const int st = q.state[row];
        rate[i] = channel_rate(q.model, q.par, st, (double)ca, q.mg);
    }
    __syncthreads();
    // the total in channel order on one lane, then every channel's share of it
    // on all lanes (the serial sums below then only add)
    if (lane == 0) {
        double tot = 0.0;
        for (int i = 0; i < n; ++i) tot += rate[i];
        s_total = tot;
    }
    __syncthreads();
    const double total = s_total;
    if (total > 0.0)
        for (int i = lane; i < n; i += 64) wv[i] = rate[i] / total;
    __syncthreads();
    double *st = q.stats + (size_t)s * QS_NSTAT;
    // lane 0: the shares' running sums in channel order (pcv[i], the sum of
    // the shares before channel i), the spread test and the event's draws;
    // then every lane tests its channels against the draw and a ballot finds
    // the first one, the channel the serial scan would stop at
    int go = 0;
    double u1 = 0.0, r = 0.0;
    uint32_t S6[6];
    if (lane == 0) {
        if (!(total > 0.0)) {
            // nothing can happen any more: N stays until t_max, and the open
            // channels keep drawing on the jSR until then
            st[QS_WQ] += st[QS_NCUR] * (q.t_max - fmin(st[QS_TCUR], q.t_max));
            if (q.deplete && q.t[s] < q.t_max) {
                const double c0 = q.jsr[s];
                double geff = 0.0;
                for (int i = 0; i < n; ++i)
                    if (q.state[off + i] == 1) geff += q.gch * (1.0 - (double)cav[i] / c0);
                const double cw = jsr_advance(q, c0, geff, q.t_max - q.t[s]);
                q.jsr[s] = cw;
                if (cw < q.jsr_min[s]) q.jsr_min[s] = cw;
            }
            q.active[s] = 0;
            q.waiting[s] = 0;
        } else {
            double pc = 0.0, m = 0.0;
            for (int i = 0; i < n; ++i) {
                pcv[i] = pc;
                m += wv[i] * (double)spv[i];
                pc += wv[i];
            }
            if (wait == 0 && m > (double)q.thresh) {
                q.waiting[s] = 1;
            } else {
                q.waiting[s] = 0;
                // Gillespie step (QuasyStationaryReleaseSite.cpp:101-128)
                for (int k = 0; k < 6; ++k) S6[k] = q.rng[s * 6 + k];
                u1 = mrg::draw(S6);
                r = mrg::draw(S6);
                go = 1;
            }
        }
    }
    go = __shfl(go, 0);
    if (!go) return;  // the whole wave
    r = __shfl(r, 0);
    __syncthreads();
    int idx = 0;
    {
        const int i0 = lane, i1 = 64 + lane;
        const bool h0 = i0 < n && pcv[i0] <= r && r <= pcv[i0] + wv[i0];
        const bool h1 = i1 < n && pcv[i1] <= r && r <= pcv[i1] + wv[i1];
        const unsigned long long b0 = __ballot(h0), b1 = __ballot(h1);
        if (b0) idx = __ffsll((long long)b0) - 1;
        else if (b1) idx = 64 + __ffsll((long long)b1) - 1;
    }
    if (lane != 0) return;
    const double dtw = -log(u1) * (1.0 / total);
    const double t = q.t[s] + dtw;
    (void)mrg::draw(S6);  // MakeTransition(runif())
    if (q.deplete) {
        // the jSR over the interval before the event, with this interval's open set
        const double c0 = q.jsr[s];
        double geff = 0.0;
        for (int i = 0; i < n; ++i)
            if (q.state[off + i] == 1) geff += q.gch * (1.0 - (double)cav[i] / c0);
        const double c1 = jsr_advance(q, c0, geff, dtw);
        q.jsr[s] = c1;
        // lowest jSR over [0, t_max], what a run of that length observes
        if (t <= q.t_max) {
            if (c1 < q.jsr_min[s]) q.jsr_min[s] = c1;
        } else if (q.t[s] < q.t_max) {
            const double cw = jsr_advance(q, c0, geff, q.t_max - q.t[s]);
            if (cw < q.jsr_min[s]) q.jsr_min[s] = cw;
        }
    }
    if (q.trace) {
        // IndChannelCREAnalyzer.cpp:78-84: bins up to the next record's time
        // take the current count
        const long long nxt = (long long)(t / q.bin_ms);
        const long long end = nxt < q.bins - 1 ? nxt : q.bins - 1;
        const uint8_t v = (uint8_t)(st[QS_NCUR] < 254.0 ? st[QS_NCUR] : 254.0);
        for (long long k = q.cur_bin[s]; k <= end; ++k) q.trace[(size_t)k * q.S + s] = v;
        q.cur_bin[s] = (int)(nxt + 1 < (long long)q.bins ? nxt + 1 : q.bins);
    }
    const int was = q.state[off + idx];
    q.state[off + idx] = was ^ 1;
    const double ncur = st[QS_NCUR];
    const double nnew = ncur + (was ? -1.0 : 1.0);
    st[QS_Q] += ncur * (t - st[QS_TCUR]);
    if (nnew > st[QS_AMP]) {
        st[QS_AMP] = nnew;
        st[QS_TTP] = t;
    }
    // the window [0, t_max]
    st[QS_WQ] += ncur * (fmin(t, q.t_max) - fmin(st[QS_TCUR], q.t_max));
    if (t <= q.t_max) {
        st[QS_WTTE] = t;
        st[QS_WN] = nnew;
        if (nnew > st[QS_WAMP]) st[QS_WAMP] = nnew;
    }
    st[QS_TCUR] = t;
    st[QS_TTE] = t;
    st[QS_NCUR] = nnew;
    st[QS_EVENTS] += 1.0;
    q.t[s] = t;
    for (int k = 0; k < 6; ++k) q.rng[s * 6 + k] = S6[k];
    if (nnew < 1.0 || t >= q.t_max) q.active[s] = 0;
}

// InitOpening of every site (QuasyStationaryReleaseSite.cpp:91-97).
__global__ void k_qs_init(QsDev q, const int *__restrict__ init, const uint32_t *__restrict__ rng0,
                          const uint32_t *__restrict__ jump, const double *__restrict__ jsr0) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= q.S) return;
    const int off = q.off[s], n = q.off[s + 1] - off;
    for (int k = 0; k < n; ++k) q.state[off + k] = (k == init[s]) ? 1 : 0;
    uint32_t S6[6], o6[6];
    for (int k = 0; k < 6; ++k) S6[k] = rng0[s * 6 + k];
    const uint32_t *J = jump + (size_t)(n + 1) * 18;
    mrg::matvec(J, S6, o6, mrg::C1, mrg::M1);
    mrg::matvec(J + 9, S6 + 3, o6 + 3, mrg::C2, mrg::M2);
    for (int k = 0; k < 6; ++k) q.rng[s * 6 + k] = o6[k];
    q.t[s] = 0.0;
    q.cur_bin[s] = 0;
    q.jsr[s] = jsr0[s];
    q.jsr_min[s] = jsr0[s];
    q.active[s] = q.t_max > 0.0;
    q.waiting[s] = 0;
    double *st = q.stats + (size_t)s * QS_NSTAT;
    for (int k = 0; k < QS_NSTAT; ++k) st[k] = 0.0;
    st[QS_AMP] = 1.0;
    st[QS_NCUR] = 1.0;
    st[QS_WAMP] = 1.0;
    st[QS_WN] = 1.0;
}

// Rows of this round: every site with `all`, else the active sites that do
// not wait for the full solver, packed in site order (one block).  cnt[0..2]:
// active sites, waiting sites, rows.
__global__ __launch_bounds__(1024) void k_qs_plan(QsDev q, int all, int *__restrict__ cnt) {
    // per-wave totals of the four counts, then their running sums (16 waves)
    __shared__ int wsum[4][16], wpre[4][16];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int per = (q.S + 1023) / 1024;
    const int s0 = t * per, s1 = min(q.S, s0 + per);
    int rows = 0, na = 0, nw = 0, nc = 0;
    for (int s = s0; s < s1; ++s) {
        const bool act = q.active[s] != 0, w = act && q.waiting[s] == 1;
        na += act;
        nw += w;
        if (all || (act && !w)) {
            rows += q.off[s + 1] - q.off[s];
            ++nc;
        }
    }
    // inclusive scans across the wave (thread order = site order)
    int v[4] = {rows, nc, na, nw};
#pragma unroll
    for (int d = 1; d < 64; d <<= 1)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int u = __shfl_up(v[k], d);
            if (lane >= d) v[k] += u;
        }
    if (lane == 63)
#pragma unroll
        for (int k = 0; k < 4; ++k) wsum[k][wv] = v[k];
    __syncthreads();
    if (t < 4) {
        int acc = 0;
        for (int j = 0; j < 16; ++j) {
            wpre[t][j] = acc;
            acc += wsum[t][j];
        }
    }
    __syncthreads();
    int r = wpre[0][wv] + v[0] - rows, c = wpre[1][wv] + v[1] - nc;
    for (int s = s0; s < s1; ++s) {
        const bool act = q.active[s] != 0, w = act && q.waiting[s] == 1;
        if (all || (act && !w)) {
            q.crow[s] = r;
            r += q.off[s + 1] - q.off[s];
            q.list[c++] = s;
        } else {
            q.crow[s] = -1;
        }
    }
    if (t == 1023) {
        cnt[0] = wpre[2][15] + v[2];
        cnt[1] = wpre[3][15] + v[3];
        cnt[2] = wpre[0][15] + v[0];
        cnt[3] = wpre[1][15] + v[1];  // sites with rows: the grid of k_qs_features and k_qs_event
    }
}

}  // namespace casim

struct CasimQS {
    int dev = 0;
    hipStream_t stream = nullptr;
    QsDev q{};
    int nch = 0;
    std::vector<int> h_off, h_init;
    std::vector<uint32_t> h_rng0;
    int *d_init = nullptr, *d_count = nullptr;
    double *d_jsr0 = nullptr;
    uint32_t *d_rng0 = nullptr, *d_jump = nullptr;
    std::vector<void *> bufs;
    void use() { HIPCK(hipSetDevice(dev)); }
    template <class T>
    T *alloc(size_t n) {
        void *p = nullptr;
        HIPCK(hipMalloc(&p, sizeof(T) * std::max<size_t>(n, 1)));
        bufs.push_back(p);
        return (T *)p;
    }
};

extern "C" {

int casim_qs_create(const CasimQSConfig *c, CasimQS **out) {
    CasimQS *h = nullptr;
    int rc = guard([&] {
        REQUIRE(c && out && c->n_sites >= 1 && c->site_channel_offset && c->channel_x && c->channel_y && c->jsr &&
                    c->rng_state && c->init_channel && c->params,
                CASIM_E_INVALID, "bad quasi-stationary configuration");
        REQUIRE(c->model >= 1 && c->model <= 3, CASIM_E_INVALID, "unknown kinetic model");
        REQUIRE(c->pitch > 0 && c->half_extent > 0, CASIM_E_INVALID, "pitch and extent must be > 0");
        if (c->jsr_deplete) {
            REQUIRE(c->jsr_volume > 0 && c->element_volume > 0 && c->jsr_tau_refill > 0 && c->conductance >= 0 &&
                        c->jsr_cq_total >= 0 && c->jsr_k_cq >= 0 && c->jsr_cq_sites >= 0,
                    CASIM_E_INVALID, "bad jSR pool parameters");
            for (int k = 0; k < c->n_sites; ++k)
                REQUIRE(c->jsr[k] > 0, CASIM_E_INVALID, "a depleting jSR must start above 0");
        }
        REQUIRE(c->t_max >= 0 && c->spread_threshold >= 0, CASIM_E_INVALID, "t_max and the threshold must be >= 0");
        *out = nullptr;
        h = new CasimQS();
        const int S = c->n_sites;
        h->h_off.assign(c->site_channel_offset, c->site_channel_offset + S + 1);
        REQUIRE(h->h_off[0] == 0, CASIM_E_INVALID, "site_channel_offset[0] must be 0");
        for (int s = 0; s < S; ++s) {
            const int n = h->h_off[s + 1] - h->h_off[s];
            REQUIRE(n >= 1 && n <= MAX_SITE_CHANNELS, CASIM_E_INVALID, "a site needs 1..127 channels");
            REQUIRE(c->init_channel[s] >= 0 && c->init_channel[s] < n, CASIM_E_RANGE, "init channel out of range");
        }
        h->nch = h->h_off[S];
        std::vector<float> x(h->nch), y(h->nch);
        for (int k = 0; k < h->nch; ++k) {
            x[k] = (float)((c->channel_x[k] - c->half_extent) / c->pitch);
            y[k] = (float)((c->channel_y[k] - c->half_extent) / c->pitch);
        }
        h->h_rng0.resize((size_t)S * 6);
        for (size_t i = 0; i < h->h_rng0.size(); ++i) {
            REQUIRE(c->rng_state[i] < (i % 6 < 3 ? mrg::M1 : mrg::M2), CASIM_E_INVALID, "bad RNG state");
            h->h_rng0[i] = (uint32_t)c->rng_state[i];
        }
        for (int k = 0; k < S; ++k) {
            const uint64_t *r = c->rng_state + (size_t)k * 6;
            REQUIRE((r[0] | r[1] | r[2]) != 0 && (r[3] | r[4] | r[5]) != 0, CASIM_E_INVALID,
                    "an RNG state has a zero component triple (RngStream::CheckSeed)");
        }
        h->h_init.assign(c->init_channel, c->init_channel + S);
        // host-side checks done: now the device
        HIPCK(hipGetDevice(&h->dev));
        HIPCK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        // jump table A^(2p), as the dyad handle builds it
        std::vector<uint32_t> jump((size_t)JUMP_N * 18);
        {
            uint32_t A1[9], A2[9], S1[9], S2[9], P1[9], P2[9];
            mrg::base_matrices(A1, A2);
            mrg::matmul(A1, A1, S1, mrg::C1, mrg::M1);
            mrg::matmul(A2, A2, S2, mrg::C2, mrg::M2);
            mrg::identity(P1);
            mrg::identity(P2);
            for (int p = 0; p < JUMP_N; ++p) {
                std::copy(P1, P1 + 9, &jump[(size_t)p * 18]);
                std::copy(P2, P2 + 9, &jump[(size_t)p * 18 + 9]);
                mrg::matmul(S1, P1, P1, mrg::C1, mrg::M1);
                mrg::matmul(S2, P2, P2, mrg::C2, mrg::M2);
            }
        }
        QsDev &q = h->q;
        q.S = S;
        int *off = h->alloc<int>(S + 1);
        float *dx = h->alloc<float>(h->nch), *dy = h->alloc<float>(h->nch);
        double *jsr = h->alloc<double>(S);
        h->d_jsr0 = h->alloc<double>(S);
        q.jsr_min = h->alloc<double>(S);
        HIPCK(hipMemcpyAsync(off, h->h_off.data(), sizeof(int) * (S + 1), hipMemcpyHostToDevice, h->stream));
        HIPCK(hipMemcpyAsync(dx, x.data(), sizeof(float) * h->nch, hipMemcpyHostToDevice, h->stream));
        HIPCK(hipMemcpyAsync(dy, y.data(), sizeof(float) * h->nch, hipMemcpyHostToDevice, h->stream));
        HIPCK(hipMemcpyAsync(jsr, c->jsr, sizeof(double) * S, hipMemcpyHostToDevice, h->stream));
        HIPCK(hipMemcpyAsync(h->d_jsr0, c->jsr, sizeof(double) * S, hipMemcpyHostToDevice, h->stream));
        q.off = off;
        q.x = dx;
        q.y = dy;
        q.jsr = jsr;
        q.half = (float)(c->half_extent / c->pitch);
        q.state = h->alloc<int>(h->nch);
        q.rng = h->alloc<uint32_t>((size_t)S * 6);
        q.t = h->alloc<double>(S);
        q.active = h->alloc<int>(S);
        q.waiting = h->alloc<int>(S);
        q.stats = h->alloc<double>((size_t)S * QS_NSTAT);
        q.X = h->alloc<float>((size_t)h->nch * QS_FMAX);
        q.nf = QS_F;
        q.lin_tab = nullptr;
        q.lin_n = 0;
        q.lin_residual = 0;
        q.ca = h->alloc<float>(h->nch);
        q.spread = h->alloc<float>(h->nch);
        {
            double *xd = h->alloc<double>(h->nch), *yd = h->alloc<double>(h->nch);
            HIPCK(hipMemcpyAsync(xd, c->channel_x, sizeof(double) * h->nch, hipMemcpyHostToDevice, h->stream));
            HIPCK(hipMemcpyAsync(yd, c->channel_y, sizeof(double) * h->nch, hipMemcpyHostToDevice, h->stream));
            q.xd = xd;
            q.yd = yd;
            q.st_n = 0;
        }
        q.ovr = h->alloc<float>(h->nch);
        q.out = nullptr;
        q.model = c->model;
        for (int k = 0; k < NPAR; ++k) q.par[k] = c->params[k];
        q.t_max = c->t_max;
        q.mg = c->mg;
        q.deplete = c->jsr_deplete != 0;
        q.Vr = q.deplete ? c->element_volume / c->jsr_volume : 0.0;
        q.cq = c->jsr_cq_total;
        q.ncq = c->jsr_cq_sites;
        q.kcq = c->jsr_k_cq;
        q.tau = q.deplete ? c->jsr_tau_refill : 1.0;
        q.nsr = c->nsr;
        q.gch = c->conductance;
        q.thresh = (float)c->spread_threshold;
        h->d_init = h->alloc<int>(S);
        h->d_count = h->alloc<int>(4);
        q.cur_bin = h->alloc<int>(S);
        q.trace = nullptr;
        q.bins = 0;
        q.bin_ms = 0.01;
        q.crow = h->alloc<int>(S);
        q.list = h->alloc<int>(S);
        h->d_rng0 = h->alloc<uint32_t>((size_t)S * 6);
        h->d_jump = h->alloc<uint32_t>(jump.size());
        HIPCK(hipMemcpyAsync(h->d_init, h->h_init.data(), sizeof(int) * S, hipMemcpyHostToDevice, h->stream));
        HIPCK(hipMemcpyAsync(h->d_rng0, h->h_rng0.data(), sizeof(uint32_t) * S * 6, hipMemcpyHostToDevice, h->stream));
        HIPCK(hipMemcpyAsync(h->d_jump, jump.data(), sizeof(uint32_t) * jump.size(), hipMemcpyHostToDevice, h->stream));
        HIPCK(hipMemsetAsync(q.ca, 0, sizeof(float) * h->nch, h->stream));
        HIPCK(hipMemsetAsync(q.spread, 0, sizeof(float) * h->nch, h->stream));
        HIPCK(hipStreamSynchronize(h->stream));
        *out = h;
        if (casim_qs_reset(h) != CASIM_OK) throw Err(CASIM_E_DEVICE, last_error());
    });
    if (rc != CASIM_OK && h) {
        std::string keep = last_error();
        casim_qs_destroy(h);
        last_error() = keep;
        if (out) *out = nullptr;
    }
    return rc;
}

int casim_qs_destroy(CasimQS *h) {
    if (!h) return CASIM_OK;
    (void)hipSetDevice(h->dev);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (void *p : h->bufs) (void)hipFree(p);
    if (h->q.out) (void)hipFree(h->q.out);
    if (h->q.trace) (void)hipFree(h->q.trace);
    if (h->q.lin_tab) (void)hipFree((void *)h->q.lin_tab);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return CASIM_OK;
}

int casim_qs_reset(CasimQS *h) {
    return guard([&] {
        REQUIRE(h, CASIM_E_INVALID, "null handle");
        h->use();
        if (h->q.trace) HIPCK(hipMemsetAsync(h->q.trace, 0xFF, (size_t)h->q.bins * h->q.S, h->stream));
        hipLaunchKernelGGL(k_qs_init, dim3((h->q.S + 127) / 128), dim3(128), 0, h->stream, h->q, h->d_init, h->d_rng0,
                           h->d_jump, (const double *)h->d_jsr0);
        HIPCK(hipGetLastError());
        HIPCK(hipStreamSynchronize(h->stream));
    });
}

int casim_qs_run(CasimQS *h, CasimMLP *m, int32_t max_rounds, int32_t wait_batch, int32_t *rounds,
                 int32_t *n_active, int32_t *n_waiting) {
    return guard([&] {
        REQUIRE(h, CASIM_E_INVALID, "null handle");
        const bool stationary = h->q.st_n > 0;
        REQUIRE(m || stationary, CASIM_E_INVALID, "null ensemble (and no stationary model set)");
        REQUIRE(stationary || mlp_inputs(m) == h->q.nf, CASIM_E_INVALID,
                h->q.lin_n ? "with a superposition table the surrogate must take 57 features"
                           : "the surrogate must take 56 features (57 with a superposition table)");
        h->use();
        if (!stationary) {
            const int E = mlp_members(m);
            if (!h->q.out || h->q.members != E) {
                if (h->q.out) (void)hipFree(h->q.out);
                h->q.out = nullptr;
                HIPCK(hipMalloc(&h->q.out, sizeof(float) * (size_t)E * h->nch));
            }
            h->q.members = E;
            h->q.ym = mlp_out_mean(m);
            h->q.ys = mlp_out_std(m);
        }
        const int batch = std::max(1, (int)wait_batch);
        int cnt[4] = {0, 0, 0, 0};
        auto plan = [&] {
            hipLaunchKernelGGL(k_qs_plan, dim3(1), dim3(1024), 0, h->stream, h->q, 0, h->d_count);
            HIPCK(hipGetLastError());
            HIPCK(hipMemcpyAsync(cnt, h->d_count, sizeof(int) * 4, hipMemcpyDeviceToHost, h->stream));
            HIPCK(hipStreamSynchronize(h->stream));
        };
        plan();
        int r = 0;
        while (r < max_rounds && cnt[0] > 0 && cnt[1] < batch && cnt[1] < cnt[0]) {
            h->q.R = cnt[2];
            if (!stationary) {
                hipLaunchKernelGGL(k_qs_features, dim3(cnt[3]), dim3(64 * QS_FWAVES), 0, h->stream, h->q);
                HIPCK(hipGetLastError());
                mlp_forward_device(m, h->q.X, cnt[2], h->q.out, h->stream);
            }
            hipLaunchKernelGGL(k_qs_event, dim3(cnt[3]), dim3(64), 0, h->stream, h->q);
            HIPCK(hipGetLastError());
            ++r;
            plan();
        }
        if (rounds) *rounds = r;
        if (n_active) *n_active = cnt[0];
        if (n_waiting) *n_waiting = cnt[1];
    });
}

int casim_qs_get_waiting(CasimQS *h, int32_t *flags, int32_t *states) {
    return guard([&] {
        REQUIRE(h, CASIM_E_INVALID, "null handle");
        h->use();
        const int S = h->q.S;
        std::vector<int> w(S), a(S), st(std::max(1, h->nch));
        HIPCK(hipMemcpyAsync(w.data(), h->q.waiting, sizeof(int) * S, hipMemcpyDeviceToHost, h->stream));
        HIPCK(hipMemcpyAsync(a.data(), h->q.active, sizeof(int) * S, hipMemcpyDeviceToHost, h->stream));
        HIPCK(hipMemcpyAsync(st.data(), h->q.state, sizeof(int) * h->nch, hipMemcpyDeviceToHost, h->stream));
        HIPCK(hipStreamSynchronize(h->stream));
        int M = 0;
        for (int s = 0; s < S; ++s) M = std::max(M, h->h_off[s + 1] - h->h_off[s]);
        if (flags)
            for (int s = 0; s < S; ++s) flags[s] = (a[s] && w[s] == 1) ? 1 : 0;
        if (states)
            for (int s = 0; s < S; ++s) {
                const int o = h->h_off[s], n = h->h_off[s + 1] - o;
                for (int j = 0; j < M; ++j) states[(size_t)s * M + j] = j < n ? st[o + j] : -1;
            }
    });
}

int casim_qs_set_mouths(CasimQS *h, int32_t site, const float *ca) {
    return guard([&] {
        REQUIRE(h && ca, CASIM_E_INVALID, "null argument");
        REQUIRE(site >= 0 && site < h->q.S, CASIM_E_RANGE, "site out of range");
        h->use();
        const int o = h->h_off[site], n = h->h_off[site + 1] - o;
        const int two = 2;
        HIPCK(hipMemcpyAsync(h->q.ovr + o, ca, sizeof(float) * n, hipMemcpyHostToDevice, h->stream));
        HIPCK(hipMemcpyAsync(h->q.waiting + site, &two, sizeof(int), hipMemcpyHostToDevice, h->stream));
        HIPCK(hipStreamSynchronize(h->stream));
    });
}

int casim_qs_get_stats(CasimQS *h, double *out) {
    return guard([&] {
        REQUIRE(h && out, CASIM_E_INVALID, "null argument");
        h->use();
        const int S = h->q.S;
        std::vector<double> st((size_t)S * QS_NSTAT);
        HIPCK(hipMemcpyAsync(st.data(), h->q.stats, sizeof(double) * st.size(), hipMemcpyDeviceToHost, h->stream));
        HIPCK(hipStreamSynchronize(h->stream));
        for (int s = 0; s < S; ++s) {
            const double *a = &st[(size_t)s * QS_NSTAT];
            double *o = out + (size_t)s * 6;
            o[0] = a[QS_AMP];
            o[1] = a[QS_TTP];
            o[2] = a[QS_TTE];
            o[3] = a[QS_Q];
            o[4] = a[QS_TTE] > 0 ? a[QS_Q] / a[QS_TTE] : 0.0;
            o[5] = a[QS_EVENTS];
        }
    });
}

int casim_qs_set_linear_feature(CasimQS *h, const float *added, int32_t n, float background, float jsr_ref,
                                int32_t residual) {
    return guard([&] {
        REQUIRE(h, CASIM_E_INVALID, "null handle");
        REQUIRE(n == 0 || (added && n >= 2 && background > 0 && jsr_ref > 0), CASIM_E_INVALID,
                "a superposition table needs >= 2 entries, background > 0 and jsr_ref > 0");
        h->use();
        HIPCK(hipStreamSynchronize(h->stream));
        if (h->q.lin_tab) (void)hipFree((void *)h->q.lin_tab);
        h->q.lin_tab = nullptr;
        h->q.lin_n = 0;
        h->q.lin_residual = 0;
        h->q.nf = QS_F;
        if (n > 0) {
            float *t = nullptr;
            HIPCK(hipMalloc(&t, sizeof(float) * n));
            HIPCK(hipMemcpyAsync(t, added, sizeof(float) * n, hipMemcpyHostToDevice, h->stream));
            HIPCK(hipStreamSynchronize(h->stream));
            h->q.lin_tab = t;
            h->q.lin_n = n;
            h->q.lin_bg = background;
            h->q.lin_jsr = jsr_ref;
            h->q.lin_residual = residual != 0;
            h->q.nf = QS_F + 1;
        }
    });
}

int casim_qs_set_stationary_model(CasimQS *h, int32_t n_modes, const double *coef, const double *decay,
                                  double background, double ca_open) {
    return guard([&] {
        REQUIRE(h, CASIM_E_INVALID, "null handle");
        REQUIRE(n_modes >= 0 && n_modes <= QS_MAXMODES, CASIM_E_RANGE,
                "a stationary model takes 0..8 modes (up to 7 buffers)");
        REQUIRE(n_modes == 0 || (coef && decay), CASIM_E_INVALID, "null mode arrays");
        REQUIRE(n_modes == 0 || !h->q.deplete, CASIM_E_INVALID,
                "the stationary model does not follow the jSR: create the sites without jsr_deplete");
        for (int k = 0; k < n_modes; ++k)
            REQUIRE(std::isfinite(coef[k]) && std::isfinite(decay[k]) && decay[k] >= 0, CASIM_E_INVALID,
                    "mode coefficients must be finite and decays >= 0");
        REQUIRE(std::isfinite(background) && std::isfinite(ca_open), CASIM_E_INVALID, "non-finite concentration");
        h->use();
        HIPCK(hipStreamSynchronize(h->stream));
        h->q.st_n = n_modes;
        for (int k = 0; k < QS_MAXMODES; ++k) {
            h->q.st_coef[k] = k < n_modes ? coef[k] : 0.0;
            h->q.st_decay[k] = k < n_modes ? decay[k] : 0.0;
        }
        h->q.st_bg = background;
        h->q.st_open = ca_open;
    });
}

int casim_qs_record_open_counts(CasimQS *h, double bin_ms, int32_t max_bins) {
    return guard([&] {
        REQUIRE(h, CASIM_E_INVALID, "null handle");
        REQUIRE(bin_ms > 0 && max_bins >= 0, CASIM_E_INVALID, "bin_ms > 0, max_bins >= 0");
        h->use();
        HIPCK(hipStreamSynchronize(h->stream));
        if (h->q.trace) (void)hipFree(h->q.trace);
        h->q.trace = nullptr;
        h->q.bins = max_bins;
        h->q.bin_ms = bin_ms;
        if (max_bins > 0) {
            HIPCK(hipMalloc(&h->q.trace, (size_t)max_bins * h->q.S));
            HIPCK(hipMemsetAsync(h->q.trace, 0xFF, (size_t)max_bins * h->q.S, h->stream));
            HIPCK(hipStreamSynchronize(h->stream));
        }
    });
}

int casim_qs_get_open_counts(CasimQS *h, uint8_t *out) {
    return guard([&] {
        REQUIRE(h && out, CASIM_E_INVALID, "null argument");
        REQUIRE(h->q.trace, CASIM_E_STATE, "open-count recording is off");
        h->use();
        HIPCK(hipMemcpyAsync(out, h->q.trace, (size_t)h->q.bins * h->q.S, hipMemcpyDeviceToHost, h->stream));
        HIPCK(hipStreamSynchronize(h->stream));
    });
}

int casim_qs_get_window_stats(CasimQS *h, double *out) {
    return guard([&] {
        REQUIRE(h && out, CASIM_E_INVALID, "null argument");
        h->use();
        const int S = h->q.S;
        std::vector<double> st((size_t)S * QS_NSTAT);
        HIPCK(hipMemcpyAsync(st.data(), h->q.stats, sizeof(double) * st.size(), hipMemcpyDeviceToHost, h->stream));
        HIPCK(hipStreamSynchronize(h->stream));
        for (int s = 0; s < S; ++s) {
            const double *a = &st[(size_t)s * QS_NSTAT];
            double *o = out + (size_t)s * 4;
            o[0] = a[QS_WAMP];
            o[1] = a[QS_WTTE];
            o[2] = a[QS_WQ];
            o[3] = a[QS_WN];
        }
    });
}

int casim_qs_get_features(CasimQS *h, float *out) {
    return guard([&] {
        REQUIRE(h && out, CASIM_E_INVALID, "null argument");
        h->use();
        hipLaunchKernelGGL(k_qs_plan, dim3(1), dim3(1024), 0, h->stream, h->q, 1, h->d_count);
        HIPCK(hipGetLastError());
        hipLaunchKernelGGL(k_qs_features, dim3(h->q.S), dim3(64 * QS_FWAVES), 0, h->stream, h->q);
        HIPCK(hipGetLastError());
        HIPCK(hipMemcpyAsync(out, h->q.X, sizeof(float) * (size_t)h->nch * h->q.nf, hipMemcpyDeviceToHost,
                             h->stream));
        HIPCK(hipStreamSynchronize(h->stream));
    });
}

int casim_qs_get_jsr(CasimQS *h, double *now, double *lowest) {
    return guard([&] {
        REQUIRE(h, CASIM_E_INVALID, "null handle");
        h->use();
        if (now) HIPCK(hipMemcpyAsync(now, h->q.jsr, sizeof(double) * h->q.S, hipMemcpyDeviceToHost, h->stream));
        if (lowest)
            HIPCK(hipMemcpyAsync(lowest, h->q.jsr_min, sizeof(double) * h->q.S, hipMemcpyDeviceToHost, h->stream));
        HIPCK(hipStreamSynchronize(h->stream));
    });
}

int casim_qs_get_mouths(CasimQS *h, float *mean, float *spread) {
    return guard([&] {
        REQUIRE(h, CASIM_E_INVALID, "null handle");
        h->use();
        if (mean) HIPCK(hipMemcpyAsync(mean, h->q.ca, sizeof(float) * h->nch, hipMemcpyDeviceToHost, h->stream));
        if (spread)
            HIPCK(hipMemcpyAsync(spread, h->q.spread, sizeof(float) * h->nch, hipMemcpyDeviceToHost, h->stream));
        HIPCK(hipStreamSynchronize(h->stream));
    });
}

}  // extern "C"
