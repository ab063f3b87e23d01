// dyad_internal.hpp -- the handle behind the C ABI and the error helpers,
// shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/casim_dyad.h"
#include "dyad_kernels.hpp"

namespace casim {

struct Err : std::runtime_error {
    int code;
    Err(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// Message of the last failed ABI call on this thread (defined in casim_dyad.hip).
std::string &last_error();

// Run an ABI body, turning exceptions into status codes.
template <class F>
int guard(F &&f) {
    try {
        f();
        last_error().clear();
        return CASIM_OK;
    } catch (Err &e) {
        last_error() = e.what();
        return e.code;
    } catch (std::exception &e) {
        last_error() = e.what();
        return CASIM_E_INVALID;
    }
}

}  // namespace casim

#define HIPCK(x)                                                                                  \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) throw casim::Err(CASIM_E_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define REQUIRE(cond, code, msg)               \
    do {                                       \
        if (!(cond)) throw casim::Err(code, msg); \
    } while (0)

namespace casim {
// Equilibrium bound concentrations of one buffer over the ions it binds
// (GaussElimination of DyadRDCPU.cpp:163-193 on the system of :53-81).
std::vector<double> buffer_equilibrium(const std::vector<double> &kon, const std::vector<double> &koff,
                                       const std::vector<double> &c, double total);
struct CytoBox;  // 3D cytosol box around each site (cytosol_box.hip)
}  // namespace casim

struct CasimDyad {
    int dev = 0;
    hipStream_t stream = nullptr;
    casim::Params P{};
    casim::Channels Ch{};
    casim::Pools Q{};           // compartments around the dyad (Q.on = 0 when not attached)
    double *d_jc0 = nullptr;    // jSR start concentration per ion
    std::vector<double *> pool_buffers;
    int precision = CASIM_FP64;
    size_t tsize = 8;
    void *g[2] = {nullptr, nullptr};
    void *f[2] = {nullptr, nullptr};
    int cur = 0;
    int nch = 0, M = 0;
    double zext = 0, dx = 0, dy = 0, dz = 0, radius = 0, voltage = 0;
    double elem_volume = 0, grad_mult = 0, sxm = 0, sym = 0;
    std::vector<int> h_off, h_node, h_perm, h_snode, h_sperm;
    std::vector<uint32_t> h_rng0;
    int *d_off = nullptr, *d_node = nullptr, *d_type = nullptr, *d_perm = nullptr, *d_snode = nullptr, *d_sperm = nullptr,
        *d_state = nullptr, *d_chans = nullptr;
    uint32_t *d_rng = nullptr, *d_rng0 = nullptr, *d_jump = nullptr;
    double *d_mouths = nullptr, *d_jsr = nullptr, *d_sJ = nullptr, *d_snap = nullptr, *d_pg = nullptr, *d_pjv = nullptr, *d_totals = nullptr, *d_rim_ions = nullptr,
           *d_rim_bound = nullptr, *d_eff = nullptr, *d_stats = nullptr;
    bool have_update = false;
    bool start_equilibrium = false;
    int threads = 256;        // k_step block size
    int sweep_threads = 0;    // k_sweep block size (0: not usable)
    int sweep_chunk = 0;      // interior x-planes per k_sweep workgroup
    int sweep_prefetch = 1;   // planes prefetched beyond x+1
    bool use_sweep = true;
    int steps_per_pass = 2;   // 2: pairs of steps in one k_sweep2 pass where pair_ok()
    bool chan_small = false;  // the channel kernel's SheepCannell / one-ion instantiation applies
    casim::CytoBox *box = nullptr;  // cytosol box (null when not attached)
    // the sites and stream the pair and channel launchers work on (all sites
    // on `stream` unless casim_dyad_step splits the batch)
    int range_s0 = 0, range_n = -1;
    hipStream_t range_stream = nullptr;
    // split pipelining: the batch's two halves on `stream` and `stream2`,
    // offset by one channel call, so one half's gating overlaps the other's pass
    bool split = true;
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_off = nullptr, ev_join = nullptr;
    int range_count() const { return range_n < 0 ? P.S : range_n; }
    hipStream_t range_st() const { return range_stream ? range_stream : stream; }
    // captured K-step graphs, keyed by (k_steps, dt, state buffers in use:
    // bit 0 the dyad's, bit 1 the box's), and the buffer flips each makes
    std::map<std::tuple<int, double, int>, std::pair<hipGraphExec_t, int>> graphs;

    size_t grid_elems() const { return (size_t)P.n_grid * P.S * P.NS; }
    size_t face_elems() const { return (size_t)P.n_face * P.S * 2 * P.nFP; }

    void use() { HIPCK(hipSetDevice(dev)); }

    void launch_step(double dt);
    bool pair_ok() const;
    void launch_pair(double dt);
    void launch_channels(int mode, double dt);
    void launch_efflux();
    void launch_rim(const double *ions, const double *bound);
    void clear_graphs() {
        for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second.first);
        graphs.clear();
    }
};

namespace casim {
// State-sweep launchers, one translation unit each (sweep_inst.hip).
void sweep_launch_f64_3d(CasimDyad *d, double dt);
void sweep_launch_f64_2d(CasimDyad *d, double dt);
void sweep_launch_f32_3d(CasimDyad *d, double dt);
void sweep_launch_f32_2d(CasimDyad *d, double dt);
void sweep2_launch_f64_3d(CasimDyad *d, double dt);
void sweep2_launch_f64_2d(CasimDyad *d, double dt);
void sweep2_launch_f32_3d(CasimDyad *d, double dt);
void sweep2_launch_f32_2d(CasimDyad *d, double dt);
}  // namespace casim

namespace casim {
// Cytosol box launches (cytosol_box.hip), all on the handle's stream.
// box_efflux: dt times the rim-link effluxes of the current dyad state into
// the box's accumulator (first = overwrite); box_step: one box step of
// substeps*dt consuming the accumulator; box_rim: dyad rim nodes from the box.
void box_efflux(CasimDyad *d, bool first, double dt);
// box_efflux_pair: after a two-step pass, both steps' effluxes in one launch
// (the pass's input buffer for the first, the ring values k_sweep2 wrote for
// the second); box_ring: that ring buffer (null without a box).
void box_efflux_pair(CasimDyad *d, bool first, double dt);
void *box_ring(const CasimDyad *d);
void box_step(CasimDyad *d, double dt);
void box_rim(CasimDyad *d);
void box_reset(CasimDyad *d);
void box_free(CasimDyad *d);
int box_substeps(const CasimDyad *d);
int box_cur(const CasimDyad *d);
void box_set_cur(CasimDyad *d, int cur);
}  // namespace casim

struct CasimMLP;
namespace casim {
// Device-side ensemble forward (mlp.hip): out [members x rows], normalised units.
void mlp_forward_device(CasimMLP *m, const float *x, int64_t rows, float *out, hipStream_t stream);
int mlp_members(const CasimMLP *m);
int mlp_inputs(const CasimMLP *m);
float mlp_out_mean(const CasimMLP *m);
float mlp_out_std(const CasimMLP *m);
}  // namespace casim
