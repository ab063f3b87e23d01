// ref_shim.cpp -- C entry points over the reference's own classes, for pinning
// the oracle.  TEST INFRASTRUCTURE ONLY.
//
// Compiled by oracle/Makefile together with these unmodified reference
// sources from /root/reference/CaSim (they build with g++ from their own files):
//   rngstream.cpp, ParallelRng.cpp, TwoStateModel.cpp, SheepCannell.cpp,
//   TwoStatePoModel.cpp, TwoStateMWC.cpp, TwoStateaHTG.cpp, Point.cpp
// The result, oracle/_ref/libcasim_ref.so, is used only by tests.
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "ParallelRng.h"
#include "Point.h"
#include "SheepCannell.h"
#include "TwoStateMWC.h"
#include "TwoStateaHTG.h"
#include "rngstream.h"

extern "C" {

// Draw n RandU01 values from the k-th RngStream constructed after
// SetPackageSeed(seed); returns the stream's state afterwards.
int ref_stream_draws(const unsigned long seed[6], int k, int n, double *out, unsigned long state_out[6]) {
    if (!RngStream::SetPackageSeed(seed)) return -1;
    std::vector<RngStream> streams(k + 1);
    for (int i = 0; i < n; ++i) out[i] = streams[k].RandU01();
    streams[k].GetState(state_out);
    return 0;
}

// Start state of the k-th stream's first substream after ResetNextSubstream.
int ref_stream_substream(const unsigned long seed[6], int k, unsigned long state_out[6]) {
    if (!RngStream::SetPackageSeed(seed)) return -1;
    std::vector<RngStream> streams(k + 1);
    streams[k].ResetNextSubstream();
    streams[k].GetState(state_out);
    return 0;
}

// Advance by AdvanceState(e, c) and report the state.
int ref_stream_advance(const unsigned long seed[6], long e, long c, unsigned long state_out[6]) {
    if (!RngStream::SetPackageSeed(seed)) return -1;
    RngStream s;
    s.AdvanceState(e, c);
    s.GetState(state_out);
    return 0;
}

// ParallelRng(seed) then n runif() calls on the calling (master) thread, as
// DyadChannels does for a JSON "Seed" (DyadChannels.cpp:18-22).
int ref_parallel_rng_draws(unsigned long seed[6], int n, double *out) {
    ParallelRng rng(seed);
    for (int i = 0; i < n; ++i) out[i] = rng.runif();
    return 0;
}

// GetRate of a two-state model in macro state `state` at mouth Ca (and Mg).
double ref_sheep_cannell_rate(const double *p, int state, double ca) {
    std::map<std::string, double> params{{"kopen0", p[0]}, {"ncaopen", p[1]}, {"kopenmax", p[2]},
                                         {"kclose0", p[3]}, {"ncaclose", p[4]}};
    SheepCannell m(params);
    std::unordered_map<std::string, double> ions{{"Calcium", ca}};
    m.SetMacroState(state, 0.0, ions);
    return m.GetRate(ions);
}

double ref_mwc_rate(const double *p, int state, double ca, double mg) {
    std::map<std::string, double> params{{"fCa", p[0]}, {"KCa", p[1]}, {"fMg", p[2]}, {"KMg", p[3]},
                                         {"KO0", p[4]}, {"KMgI", p[5]}, {"kclose", p[6]}};
    TwoStateMWC m(params);
    std::unordered_map<std::string, double> ions{{"Calcium", ca}, {"Magnesium", mg}};
    m.SetMacroState(state, 0.0, ions);
    return m.GetRate(ions);
}

double ref_ahtg_rate(const double *p, int state, double ca, double mg) {
    std::map<std::string, double> params{{"fCa", p[0]}, {"KCa", p[1]}, {"fMg", p[2]}, {"KMg", p[3]},
                                         {"KO0", p[4]}, {"KMgI", p[5]}, {"KCL", p[6]}, {"KI", p[7]},
                                         {"KOL", p[8]}, {"kclose", p[9]}};
    TwoStateaHTG m(params);
    std::unordered_map<std::string, double> ions{{"Calcium", ca}, {"Magnesium", mg}};
    m.SetMacroState(state, 0.0, ions);
    return m.GetRate(ions);
}

// Point::GetCoordsOnGrid.
void ref_coords_on_grid(double x, double y, double z, double dx, double dy, double dz, int out[3]) {
    std::vector<double> c{x, y, z};
    Point p(c);
    std::vector<int> g = p.GetCoordsOnGrid(dx, dy, dz);
    out[0] = g[0];
    out[1] = g[1];
    out[2] = g[2];
}

}  // extern "C"
