// sweep_inst.hip -- launchers of the plane-marching state sweep for one
// (precision, dimension) pair, selected by CASIM_SWEEP_FP64 / CASIM_SWEEP_DIM;
// built four times so the instantiations compile in parallel.
#include "dyad_internal.hpp"

namespace casim {
namespace {

template <typename T, int DIM, int NI, int NB, int NG, int NF, int SL, int PD>
void sweep_launch(CasimDyad *d, double dt) {
    const Params &p = d->P;
    const int P = p.ny * p.nz;
    const int chunk = d->sweep_chunk;
    const int nchunks = (p.nx - 2 + chunk - 1) / chunk;
    const int units = p.S * nchunks;
    const int grid = (units + 7) / 8 * 8;
    const size_t lds = 2 * (size_t)p.n_grid * P * sizeof(T);
    hipLaunchKernelGGL((k_sweep<T, DIM, NI, NB, NG, NF, SL, PD>), dim3(grid), dim3(d->sweep_threads), lds, d->stream,
                       (const T *)d->g[d->cur], (T *)d->g[d->cur ^ 1], (const T *)d->f[d->cur],
                       (T *)d->f[d->cur ^ 1], d->d_off, d->d_snode, d->d_sJ, d->nch, d->P, dt, chunk, nchunks,
                       units);
}

template <typename T, int DIM, int NI, int NB, int NG, int NF, int SL>
void sweep_depth(CasimDyad *d, double dt) {
    if (NG == 0 || d->sweep_prefetch <= 1) return sweep_launch<T, DIM, NI, NB, NG, NF, SL, 1>(d, dt);
    if (d->sweep_prefetch == 2) return sweep_launch<T, DIM, NI, NB, NG, NF, SL, 2>(d, dt);
    return sweep_launch<T, DIM, NI, NB, NG, NF, SL, 3>(d, dt);
}

template <typename T, int DIM, int NI, int NB, int NG, int NF>
void sweep_slots(CasimDyad *d, double dt) {
    const int P = d->P.ny * d->P.nz;
    const int sl = (P + d->sweep_threads - 1) / d->sweep_threads;
    if (sl <= 1) return sweep_depth<T, DIM, NI, NB, NG, NF, 1>(d, dt);
    if (sl <= 2) return sweep_depth<T, DIM, NI, NB, NG, NF, 2>(d, dt);
    return sweep_launch<T, DIM, NI, NB, NG, NF, 4, 1>(d, dt);
}

// Species layouts with their own instantiation (ions, bound fields, grid
// fields, face fields); anything else runs the generic form.
template <typename T, int DIM>
void sweep_species(CasimDyad *d, double dt) {
    const Params &p = d->P;
    if (p.ni == 1 && p.nb == 3 && p.n_grid == 3 && p.n_face == 1) return sweep_slots<T, DIM, 1, 3, 3, 1>(d, dt);
    if (p.ni == 1 && p.nb == 3 && p.n_grid == 4 && p.n_face == 0) return sweep_slots<T, DIM, 1, 3, 4, 0>(d, dt);
    if (p.ni == 2 && p.nb == 4 && p.n_grid == 6 && p.n_face == 0) return sweep_slots<T, DIM, 2, 4, 6, 0>(d, dt);
    return sweep_slots<T, DIM, 0, 0, 0, 0>(d, dt);
}




}  // namespace

#if CASIM_SWEEP_FP64 && CASIM_SWEEP_DIM == 3
void sweep_launch_f64_3d(CasimDyad *d, double dt) { sweep_species<double, 3>(d, dt); }
#elif CASIM_SWEEP_FP64 && CASIM_SWEEP_DIM == 2
void sweep_launch_f64_2d(CasimDyad *d, double dt) { sweep_species<double, 2>(d, dt); }
#elif CASIM_SWEEP_DIM == 3
void sweep_launch_f32_3d(CasimDyad *d, double dt) { sweep_species<float, 3>(d, dt); }
#else
void sweep_launch_f32_2d(CasimDyad *d, double dt) { sweep_species<float, 2>(d, dt); }
#endif

}  // namespace casim
