// host_check.cpp -- drives libcasim_dyad.so from C++ through DyadSeriesApi,
// the way CaSim's solvers drive a dyad, and checks it on the GPU:
//   1. the reference's step protocol (BatchLoop: RunRD, RunMC per site,
//      Update) through the DyadRDHIP adapter equals the fused path
//      (casim_dyad_step, one captured graph) bit for bit: every field, the
//      channel states and every site's RNG state, with fast gating;
//   2. ABI errors arrive as exceptions carrying casim_last_error();
//   3. the output dataset shapes (GetDimensions & co.) are DyadRDCPU's, and
//      GetConcentrations returns exactly the fields asked for.
// Workload: 4 sites of a 3D cleft 80 x 80 x 16 nm at h = 4 nm (21 x 21 x 5),
// Ca2+ with mobile buffers A, B and the membrane buffer F on the z-faces,
// 3 RyRs per site with SheepCannell gating sped up so that channels toggle.
// Prints one JSON line; exit status 0 when every check holds.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "dyad_series.hpp"

using casim_host::BatchLoop;
using casim_host::DyadRDHIP;

namespace {

struct Batch {
    CasimDyadConfig c{};
    std::vector<double> ion_D{220000.0}, ion_c0{0.1}, total{5000.0, 24.0, 1000.0};
    std::vector<int32_t> faces{0, 0, 1}, bion{0, 0, 0}, bbuf{0, 1, 2};
    std::vector<double> bD{140000.0, 25000.0, 0.0}, kon{0.225, 0.1, 0.115}, koff{45.0, 0.038, 1.0};
    std::vector<int32_t> sl{0};
    std::vector<double> zero{0.0};
    std::vector<int32_t> model{CASIM_MODEL_SHEEP_CANNELL};
    std::vector<double> params = std::vector<double>(CASIM_MAX_MODEL_PARAMS, 0.0);
    std::vector<double> g{28339.0};
    std::vector<int32_t> off, type;
    std::vector<double> cx, cy;
    std::vector<uint64_t> rng;

    explicit Batch(int S) {
        // kopen = 0.2/ms ([Ca]/15 uM)^2 and kclose = 0.5/ms, both x 1e3
        params[0] = 0.2 / 225.0 * 1e9 * 1e3;
        params[1] = 2.0;
        params[2] = 1e300;
        params[3] = 500.0 * 1e3;
        params[4] = 0.0;
        const double pos[3][2] = {{40.0, 40.0}, {44.0, 40.0}, {40.0, 48.0}};
        off.push_back(0);
        for (int s = 0; s < S; ++s) {
            for (auto &p : pos) {
                cx.push_back(p[0]);
                cy.push_back(p[1]);
                type.push_back(0);
            }
            off.push_back((int32_t)cx.size());
        }
        const uint64_t seed[6] = {12345, 12345, 12345, 12345, 12345, 12345};
        rng.resize(6 * (size_t)S);
        casim_host::check(casim_rng_stream_states(seed, 0, S, rng.data()));
        c.dim = 3;
        c.nx = c.ny = 21;
        c.nz = 5;
        c.dx = c.dy = c.dz = 4.0;
        c.z = 16.0;
        c.radius = 1000.0;
        c.n_sites = S;
        c.n_ions = 1;
        c.n_buffers = 3;
        c.n_bound = 3;
        c.ion_D = ion_D.data();
        c.ion_c0 = ion_c0.data();
        c.buffer_total = total.data();
        c.buffer_faces_only = faces.data();
        c.bound_ion = bion.data();
        c.bound_buffer = bbuf.data();
        c.bound_D = bD.data();
        c.bound_kon = kon.data();
        c.bound_koff = koff.data();
        c.start_equilibrium = 1;
        c.sl_binds = sl.data();
        c.sl_N1 = c.sl_K1 = c.sl_N2 = c.sl_K2 = zero.data();
        c.calcium_index = 0;
        c.magnesium_index = -1;
        c.n_channel_types = 1;
        c.type_model = model.data();
        c.type_params = params.data();
        c.type_conductance = g.data();
        c.site_channel_offset = off.data();
        c.channel_x = cx.data();
        c.channel_y = cy.data();
        c.channel_type = type.data();
        c.rng_state = rng.data();
        c.precision = CASIM_FP64;
    }
};

}  // namespace

int main() {
    const int S = 4, steps = 60;
    const double dt = 1e-5;
    Batch b(S);
    const std::vector<std::string> ions{"Calcium"}, fields{"Calcium", "Calcium-A", "Calcium-B", "Calcium-F"};
    DyadRDHIP proto(b.c, ions, fields), fused(b.c, ions, fields);
    const int M = proto.GetNumSRChannels();

    // constant rim at the resting/equilibrium values the dyad starts from
    proto.Reset();
    std::vector<std::string> all = fields;
    auto init = proto.GetConcentrations(all);
    std::vector<double> ri(S), rb(3 * (size_t)S);
    for (int s = 0; s < S; ++s) {
        const size_t n = (size_t)s * 21 * 21 * 5;  // node (0, 0, 0) of site s: on the rim
        ri[s] = init["Calcium"][n];
        rb[s] = init["Calcium-A"][n];
        rb[S + s] = init["Calcium-B"][n];
        rb[2 * S + s] = init["Calcium-F"][n];
    }
    std::vector<double> jsr((size_t)M * S, 1000.0);
    std::vector<int> first(S, 0);

    // the reference's step protocol through the DyadSeries interface
    BatchLoop loop(proto, ri, rb, jsr);
    loop.start(first);
    loop.run(dt, steps, S);

    // the fused path on an identical batch
    fused.Reset();
    casim_host::check(casim_dyad_init_opening(fused.handle(), -1, 0, first.data()));
    double v = 0.0;
    double *pri = ri.data(), *prb = rb.data();
    std::vector<double> ext;
    fused.Update(pri, prb, jsr, ext, v);
    casim_host::check(casim_dyad_step(fused.handle(), dt, steps));
    casim_host::check(casim_dyad_synchronize(fused.handle()));

    bool fields_equal = true;
    auto a = proto.GetConcentrations(all), f = fused.GetConcentrations(all);
    for (auto &kv : a) fields_equal = fields_equal && kv.second == f[kv.first];
    std::vector<std::string> none;
    auto sa = proto.GetChannelsStates(none)["RyR"], sf = fused.GetChannelsStates(none)["RyR"];
    std::vector<uint64_t> ra(6 * (size_t)S), rf(6 * (size_t)S);
    casim_host::check(casim_dyad_get_rng_states(proto.handle(), ra.data()));
    casim_host::check(casim_dyad_get_rng_states(fused.handle(), rf.data()));
    int toggled = 0;
    for (size_t i = 0; i < sa.size(); ++i) toggled += sa[i] != (i % (size_t)M == 0 ? 1 : 0);
    bool rng_moved = ra != b.rng;

    // ABI errors surface as exceptions with the library's message
    bool error_reported = false;
    std::string msg;
    try {
        DyadRDHIP fresh(b.c, ions, fields);
        fresh.RunRD(dt, 0);  // before the first Update
    } catch (const std::runtime_error &e) {
        msg = e.what();
        error_reported = msg.find("before the first update") != std::string::npos;
    }
    // output dataset shapes the solver copies for its HDF5 writer (DyadRDCPU.cpp:21-25,137-142)
    const bool dims_ok = proto.GetDimensions() == std::vector<uint64_t>{0, (uint64_t)S, 21, 21, 5} &&
                         proto.GetChannelsDimensions() == std::vector<uint64_t>{0, (uint64_t)S, (uint64_t)M} &&
                         proto.GetIonsNearChannelsDimensions() == std::vector<uint64_t>{0, (uint64_t)S, (uint64_t)M};
    // GetConcentrations returns the fields asked for: an ion by name, bound fields by buffer name
    std::vector<std::string> some{"Calcium", "B"};
    auto sel = proto.GetConcentrations(some);
    const bool names_ok = sel.size() == 2 && sel.count("Calcium") && sel.count("Calcium-B") &&
                          sel["Calcium"].size() == (size_t)S * 21 * 21 * 5;
    const bool ok = fields_equal && sa == sf && ra == rf && rng_moved && error_reported && dims_ok && names_ok;
    std::printf("{\"fields_equal\": %s, \"states_equal\": %s, \"rng_equal\": %s, \"channels_changed\": %d, "
                "\"error_reported\": %s, \"dims_ok\": %s, \"names_ok\": %s, \"steps\": %d, \"sites\": %d, "
                "\"ok\": %s}\n",
                fields_equal ? "true" : "false", sa == sf ? "true" : "false", ra == rf ? "true" : "false", toggled,
                error_reported ? "true" : "false", dims_ok ? "true" : "false", names_ok ? "true" : "false", steps, S,
                ok ? "true" : "false");
    return ok ? 0 : 1;
}
