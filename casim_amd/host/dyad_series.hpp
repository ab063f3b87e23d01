// dyad_series.hpp -- C++ host side of libcasim_dyad.so.
//
// DyadSeriesApi has the method set and signatures of CaSim's dyad plugin
// interface (DyadSeries.h:9-44), so code written against that interface
// drives the device path unchanged; DyadRDHIP implements it over the C ABI
// (include/casim_dyad.h), and BatchLoop runs one batch with the step
// protocol of ParallelHybridSolver::RunSimulationsSRPool
// (ParallelHybridSolver.cpp:340-370): Reset, InitOpening per site, Update,
// then per step one batch-wide RunRD, RunMC for every site, Update.
//
// Errors: every ABI status other than CASIM_OK becomes a std::runtime_error
// carrying casim_last_error() (the reference itself reports nothing).
#pragma once

#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/casim_dyad.h"

namespace casim_host {

inline void check(int rc) {
    if (rc != CASIM_OK) throw std::runtime_error(std::string("casim: ") + casim_last_error());
}

// The dyad plugin interface of DyadSeries.h:9-44, method for method.
class DyadSeriesApi {
  public:
    virtual ~DyadSeriesApi() = default;
    virtual int GetNumSRChannels() = 0;
    virtual bool UsesGPU() = 0;
    virtual void Reset() = 0;
    virtual void InitOpening(int site, int channel) = 0;
    virtual double GetElementVolume() = 0;
    virtual std::vector<std::string> GetListofIons() = 0;
    virtual void GetEffluxes(double *&out) = 0;
    virtual std::vector<double> GetTotalSRCurrent() = 0;
    virtual void RunMC(double dt, int site) = 0;
    virtual void Update(double *&rim_ions, double *&rim_bound, const std::vector<double> &jsr,
                        const std::vector<double> &extracell, double &voltage) = 0;
    virtual double GetL() = 0;
    virtual void RunRD(double dt, int site) = 0;
    virtual std::map<std::string, std::vector<double>> GetConcentrations(std::vector<std::string> &names) = 0;
    virtual std::map<std::string, std::vector<int>> GetChannelsStates(std::vector<std::string> &names) = 0;
    virtual std::map<std::string, std::vector<double>> GetIonsNearChannels(std::vector<std::string> &names) = 0;
    // Output dataset shapes (DyadSeries.h:32-40): [time, site, nx, ny(, nz)] for the grid fields,
    // [time, site, channels] for channel states and mouth values; entry 0 (time) is left 0 for the
    // writer to fill, as DyadRDCPU's constructor leaves it (DyadRDCPU.cpp:21-25,137-142).
    std::vector<uint64_t> GetDimensions() { return dyad_dims; }
    std::vector<uint64_t> GetChannelsDimensions() { return channels_dims; }
    std::vector<uint64_t> GetIonsNearChannelsDimensions() { return channels_ions_dims; }

  protected:
    std::vector<uint64_t> dyad_dims, channels_ions_dims, channels_dims;
};

// A batch of sites on the GPU behind DyadSeriesApi.  The configuration is the
// flattened form of the "Dyad" JSON object (CasimDyadConfig); field_names are
// the species names in the ABI's field order (ions, then "Ion-Buffer").
class DyadRDHIP : public DyadSeriesApi {
  public:
    DyadRDHIP(const CasimDyadConfig &cfg, std::vector<std::string> ion_names, std::vector<std::string> field_names)
        : ions_(std::move(ion_names)), fields_(std::move(field_names)), S_(cfg.n_sites), ni_(cfg.n_ions),
          nb_(cfg.n_bound), nu_(cfg.n_buffers), nodes_((size_t)cfg.nx * cfg.ny * (cfg.dim == 3 ? cfg.nz : 1)),
          radius_(cfg.radius) {
        check(casim_dyad_create(&cfg, &d_));
        M_ = casim_dyad_max_channels(d_);
        // DyadRDCPU.cpp:21-25 (3D) / DyadRD2DCPU.cpp:19-22 (2D), and DyadRDCPU.cpp:137-142
        dyad_dims = {0, (uint64_t)S_, (uint64_t)cfg.nx, (uint64_t)cfg.ny};
        if (cfg.dim == 3) dyad_dims.push_back((uint64_t)cfg.nz);
        channels_ions_dims = {0, (uint64_t)S_, (uint64_t)M_};
        channels_dims = {0, (uint64_t)S_, (uint64_t)M_};
    }
    ~DyadRDHIP() override { casim_dyad_destroy(d_); }
    DyadRDHIP(const DyadRDHIP &) = delete;
    DyadRDHIP &operator=(const DyadRDHIP &) = delete;

    CasimDyad *handle() { return d_; }
    int sites() const { return S_; }

    int GetNumSRChannels() override { return M_; }
    bool UsesGPU() override { return true; }
    void Reset() override { check(casim_dyad_reset(d_)); }
    void InitOpening(int site, int channel) override { check(casim_dyad_init_opening(d_, site, channel, nullptr)); }
    double GetElementVolume() override { return casim_dyad_element_volume(d_); }
    std::vector<std::string> GetListofIons() override { return ions_; }
    double GetL() override { return radius_; }
    // batch-wide, as the reference's GPU dyads do (one call per step)
    void RunRD(double dt, int) override { check(casim_dyad_run_rd(d_, dt, -1)); }
    // the solver calls RunMC(dt, k) for every site; the device gates all at once
    void RunMC(double dt, int site) override {
        if (site == 0) check(casim_dyad_run_mc(d_, dt, -1));
    }
    void Update(double *&rim_ions, double *&rim_bound, const std::vector<double> &jsr, const std::vector<double> &,
                double &voltage) override {
        check(casim_dyad_update(d_, rim_ions, rim_bound, jsr.empty() ? nullptr : jsr.data(), voltage));
        eff_.resize((size_t)(ni_ + nb_) * S_);
        check(casim_dyad_get_effluxes(d_, eff_.data()));
    }
    void GetEffluxes(double *&out) override { out = eff_.data(); }
    std::vector<double> GetTotalSRCurrent() override {
        std::vector<double> t((size_t)ni_ * S_);
        check(casim_dyad_get_total_sr_current(d_, t.data()));
        return t;
    }
    std::map<std::string, std::vector<double>> GetConcentrations(std::vector<std::string> &names) override {
        std::map<std::string, std::vector<double>> out;
        for (int f = 0; f < ni_ + nb_; ++f) {
            const std::string &nm = fields_[f];
            const std::string buf = nm.find('-') == std::string::npos ? nm : nm.substr(nm.find('-') + 1);
            bool want = false;
            for (auto &n : names) want = want || n == nm || n == buf;
            if (!want) continue;
            std::vector<double> v((size_t)S_ * nodes_);
            check(casim_dyad_get_field(d_, f, v.data()));
            out[nm] = std::move(v);
        }
        return out;
    }
    std::map<std::string, std::vector<int>> GetChannelsStates(std::vector<std::string> &) override {
        std::vector<int> st((size_t)S_ * M_);
        check(casim_dyad_get_channel_states(d_, st.data()));
        return {{"RyR", st}};
    }
    std::map<std::string, std::vector<double>> GetIonsNearChannels(std::vector<std::string> &) override {
        std::vector<double> m((size_t)(ni_ + nb_ + nu_) * S_ * M_);
        check(casim_dyad_get_mouths(d_, m.data()));
        std::map<std::string, std::vector<double>> out;
        for (int i = 0; i < ni_; ++i)
            out[ions_[i]] = std::vector<double>(m.begin() + (size_t)i * S_ * M_, m.begin() + (size_t)(i + 1) * S_ * M_);
        return out;
    }

  private:
    CasimDyad *d_ = nullptr;
    std::vector<std::string> ions_, fields_;
    int S_, ni_, nb_, nu_, M_ = 0;
    size_t nodes_;
    double radius_;
    std::vector<double> eff_;
};

// One batch with a constant rim and jSR, in the reference's step order.
class BatchLoop {
  public:
    BatchLoop(DyadSeriesApi &dyad, std::vector<double> rim_ions, std::vector<double> rim_bound, std::vector<double> jsr)
        : dyad_(dyad), ri_(std::move(rim_ions)), rb_(std::move(rim_bound)), jsr_(std::move(jsr)) {}

    // ResetSystem, InitOpening of channels[k] on site k, Update (RunSimulationsSRPool's batch prologue)
    void start(const std::vector<int> &channels) {
        dyad_.Reset();
        for (size_t k = 0; k < channels.size(); ++k) dyad_.InitOpening((int)k, channels[k]);
        update();
    }
    // n steps of RunRD, RunMC for every site, Update
    void run(double dt, int n, int n_sites) {
        for (int j = 0; j < n; ++j) {
            dyad_.RunRD(dt, 0);
            for (int k = 0; k < n_sites; ++k) dyad_.RunMC(dt, k);
            update();
        }
    }

  private:
    void update() {
        double *ri = ri_.data(), *rb = rb_.empty() ? nullptr : rb_.data();
        double v = 0.0;
        dyad_.Update(ri, rb, jsr_, extracell_, v);
    }
    DyadSeriesApi &dyad_;
    std::vector<double> ri_, rb_, jsr_, extracell_;
};

}  // namespace casim_host
