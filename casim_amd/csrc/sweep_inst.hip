// sweep_inst.hip -- launchers of the plane-marching state sweep for one
// (precision, dimension) pair, selected by CASIM_SWEEP_FP64 / CASIM_SWEEP_DIM;
// built four times so the instantiations compile in parallel.
#include "dyad_internal.hpp"

namespace casim {
namespace {

template <typename T, int DIM, int NI, int NB, int NG, int NF, int LY, int SL, int PD>
void sweep_launch(CasimDyad *d, double dt) {
    const Params &p = d->P;
    const int P = p.ny * p.nz;
    const int chunk = d->sweep_chunk;
    const int nchunks = (p.nx - 2 + chunk - 1) / chunk;
    const int units = p.S * nchunks;
    const int grid = (units + 7) / 8 * 8;
    const size_t lds = 2 * (size_t)p.n_grid * P * sizeof(T);
    hipLaunchKernelGGL((k_sweep<T, DIM, NI, NB, NG, NF, LY, SL, PD>), dim3(grid), dim3(d->sweep_threads), lds, d->stream,
                       (const T *)d->g[d->cur], (T *)d->g[d->cur ^ 1], (const T *)d->f[d->cur],
                       (T *)d->f[d->cur ^ 1], d->d_off, d->d_snode, d->d_sJ, d->nch, d->P, dt, chunk, nchunks,
                       units);
}

template <typename T, int DIM, int NI, int NB, int NG, int NF, int LY, int SL>
void sweep_depth(CasimDyad *d, double dt) {
    if (NG == 0 || d->sweep_prefetch <= 1) return sweep_launch<T, DIM, NI, NB, NG, NF, LY, SL, 1>(d, dt);
    if (d->sweep_prefetch == 2) return sweep_launch<T, DIM, NI, NB, NG, NF, LY, SL, 2>(d, dt);
    return sweep_launch<T, DIM, NI, NB, NG, NF, LY, SL, 3>(d, dt);
}

template <typename T, int DIM, int NI, int NB, int NG, int NF, int LY>
void sweep_slots(CasimDyad *d, double dt) {
    const int P = d->P.ny * d->P.nz;
    const int sl = (P + d->sweep_threads - 1) / d->sweep_threads;
    if (sl <= 1) return sweep_depth<T, DIM, NI, NB, NG, NF, LY, 1>(d, dt);
    if (sl <= 2) return sweep_depth<T, DIM, NI, NB, NG, NF, LY, 2>(d, dt);
    return sweep_launch<T, DIM, NI, NB, NG, NF, LY, 4, 1>(d, dt);
}

// The compile-time species layout (Layout<LY>) matching the run-time one, or 0.
int layout_id(const Params &p) {
    if (p.ni != 1 || p.nb != 3 || p.nu != 3 || p.sl[0] || p.Db[0] == 0.0 || p.Db[1] == 0.0 || p.Db[2] != 0.0)
        return 0;
    for (int ly = 1; ly <= 2; ++ly) {
        bool ok = true;
        for (int b = 0; b < 3; ++b) {
            const int bion = ly == 1 ? Layout<1>::bion[b] : Layout<2>::bion[b];
            const int bbuf = ly == 1 ? Layout<1>::bbuf[b] : Layout<2>::bbuf[b];
            const int bface = ly == 1 ? Layout<1>::bface[b] : Layout<2>::bface[b];
            const int bslot = ly == 1 ? Layout<1>::bslot[b] : Layout<2>::bslot[b];
            const int uf = ly == 1 ? Layout<1>::ufaces[b] : Layout<2>::ufaces[b];
            ok = ok && p.bion[b] == bion && p.bbuf[b] == bbuf && p.bface[b] == bface && p.bslot[b] == bslot &&
                 p.ufaces[b] == uf;
        }
        if (ok && p.n_grid == (ly == 1 ? 3 : 4) && p.n_face == (ly == 1 ? 1 : 0)) return ly;
    }
    return 0;
}

template <typename T, int DIM, int NI, int NB, int NG, int NF, int LY, int SL, bool RING>
void sweep2_ring(CasimDyad *d, double dt) {
    const Params &p = d->P;
    const int P = p.ny * p.nz;
    const int chunk = d->sweep_chunk;
    const int nchunks = (p.nx - 2 + chunk - 1) / chunk;
    const int units = d->range_count() * nchunks;  // the launch range's sites
    const int grid = (units + 7) / 8 * 8;
    const size_t lds = (DIM == 3 ? 4 : 2) * (size_t)p.n_grid * P * sizeof(T);  // k_sweep2's LDS planes
    hipLaunchKernelGGL((k_sweep2<T, DIM, NI, NB, NG, NF, LY, SL, RING>), dim3(grid), dim3(d->sweep_threads), lds,
                       d->range_st(),
                       (const T *)d->g[d->cur], (T *)d->g[d->cur ^ 1], (const T *)d->f[d->cur],
                       (T *)d->f[d->cur ^ 1], d->d_off, d->d_snode, d->d_sJ,
                       (const double *)(d->Q.on ? d->Q.jnew : nullptr), (T *)box_ring(d), d->Ch, d->P, dt, chunk,
                       nchunks, units, d->range_s0);
}

// The ring store (step-k+1 values next to the rim, for a cytosol box) is a
// separate instantiation so that the plain path carries none of its code.
template <typename T, int DIM, int NI, int NB, int NG, int NF, int LY, int SL>
void sweep2_launch(CasimDyad *d, double dt) {
    if (box_ring(d)) return sweep2_ring<T, DIM, NI, NB, NG, NF, LY, SL, true>(d, dt);
    return sweep2_ring<T, DIM, NI, NB, NG, NF, LY, SL, false>(d, dt);
}

template <typename T, int DIM, int NI, int NB, int NG, int NF, int LY>
void sweep2_slots(CasimDyad *d, double dt) {
    const int P = d->P.ny * d->P.nz;
    const int sl = (P + d->sweep_threads - 1) / d->sweep_threads;
    if (sl <= 1) return sweep2_launch<T, DIM, NI, NB, NG, NF, LY, 1>(d, dt);
    if (sl <= 2) return sweep2_launch<T, DIM, NI, NB, NG, NF, LY, 2>(d, dt);
    return sweep2_launch<T, DIM, NI, NB, NG, NF, LY, 4>(d, dt);
}

template <typename T, int DIM>
void sweep2_species(CasimDyad *d, double dt) {
    const Params &p = d->P;
    const int ly = layout_id(p);
    if (DIM == 3 && ly == 1) return sweep2_slots<T, DIM, 1, 3, 3, 1, 1>(d, dt);
    if (ly == 2) return sweep2_slots<T, DIM, 1, 3, 4, 0, 2>(d, dt);
    if (p.ni == 2 && p.nb == 4 && p.n_grid == 6 && p.n_face == 0) return sweep2_slots<T, DIM, 2, 4, 6, 0, 0>(d, dt);
    return sweep2_slots<T, DIM, 0, 0, 0, 0, 0>(d, dt);
}

// Species layouts with their own instantiation (ions, bound fields, grid
// fields, face fields); anything else runs the generic form.
template <typename T, int DIM>
void sweep_species(CasimDyad *d, double dt) {
    const Params &p = d->P;
    const int ly = layout_id(p);
    if (DIM == 3 && ly == 1) return sweep_slots<T, DIM, 1, 3, 3, 1, 1>(d, dt);
    if (ly == 2) return sweep_slots<T, DIM, 1, 3, 4, 0, 2>(d, dt);
    if (p.ni == 2 && p.nb == 4 && p.n_grid == 6 && p.n_face == 0) return sweep_slots<T, DIM, 2, 4, 6, 0, 0>(d, dt);
    return sweep_slots<T, DIM, 0, 0, 0, 0, 0>(d, dt);
}




}  // namespace

#if CASIM_SWEEP_FP64 && CASIM_SWEEP_DIM == 3
void sweep_launch_f64_3d(CasimDyad *d, double dt) { sweep_species<double, 3>(d, dt); }
void sweep2_launch_f64_3d(CasimDyad *d, double dt) { sweep2_species<double, 3>(d, dt); }
#elif CASIM_SWEEP_FP64 && CASIM_SWEEP_DIM == 2
void sweep_launch_f64_2d(CasimDyad *d, double dt) { sweep_species<double, 2>(d, dt); }
void sweep2_launch_f64_2d(CasimDyad *d, double dt) { sweep2_species<double, 2>(d, dt); }
#elif CASIM_SWEEP_DIM == 3
void sweep_launch_f32_3d(CasimDyad *d, double dt) { sweep_species<float, 3>(d, dt); }
void sweep2_launch_f32_3d(CasimDyad *d, double dt) { sweep2_species<float, 3>(d, dt); }
#else
void sweep_launch_f32_2d(CasimDyad *d, double dt) { sweep_species<float, 2>(d, dt); }
void sweep2_launch_f32_2d(CasimDyad *d, double dt) { sweep2_species<float, 2>(d, dt); }
#endif

}  // namespace casim
