"""Simulation output in the reference's HDF5 layout.

ParallelHybridSolver writes one file per experiment (ParallelHybridSolver.cpp:
46-52) with the groups /Dyad /Cytosol /SR /Channels /IonsChannels, one dataset
per species, dims [time, site, ...], chunked by one time index and one batch
of sites, extended as time indices and batches arrive (_ExtendDataset and
_WriteDataPart, :177-188, :234-312; _WriteData, :316-337).  ReferenceOutput
reproduces those rules over any h5py-like file object:

  - a group's datasets are created at its first write with shape
    [1, n_threads, ...], maxshape [n_iter // it + 1, n_scenarios, ...],
    chunks [1, n_threads, ...], float64 (int32 for /Channels);
  - a later write at time index t of batch b first extends the time axis by
    one when t equals its length (and the length is below the maximum),
    otherwise the site axis by n_threads when b * n_threads equals its
    length, then writes [t, last n_threads sites];
  - writes happen at steps j with j % it == 0 for each group's stride.

Dataset contents follow the reference's getters: /Dyad the ions named in
"Outputs" and the (ion, buffer) fields of the buffers named there
(DyadRDCPU.cpp:516-539), [site, x, y(, z)]; /Channels "RyR" states
(:541-550), /IonsChannels every ion at the mouths (:552-560) and /SR the jSR
per channel (SRPool.cpp:90-101), [site, channel]; /Cytosol the cytosol pool's
ions and bound fields (CytosolPool.cpp:331-360), [site].  A cytosol box, an
extension, writes its fields to /Cytosol as [site, z, y, x].

The file object is h5py.File when h5py is importable; otherwise MemoryStore,
which implements the same subset of h5py's interface in memory and saves the
datasets to an .npz archive (keys "Group/Dataset") on close.
"""
import os

import numpy as np

GROUPS = ("Dyad", "Cytosol", "SR", "Channels", "IonsChannels")


class MemoryStore:
    """The part of h5py's File / Group / Dataset interface that
    ReferenceOutput uses, held in memory; close() saves an .npz archive."""

    class Dataset:
        def __init__(self, shape, maxshape, chunks, dtype):
            self.maxshape = tuple(maxshape)
            self.chunks = tuple(chunks)
            self.dtype = np.dtype(dtype)
            self._a = np.zeros(shape, dtype=self.dtype)

        @property
        def shape(self):
            return self._a.shape

        def resize(self, shape):
            shape = tuple(int(v) for v in shape)
            if len(shape) != len(self.maxshape) or any(s > m for s, m in zip(shape, self.maxshape)):
                raise ValueError("resize %s beyond maxshape %s" % (shape, self.maxshape))
            a = np.zeros(shape, dtype=self.dtype)  # HDF5's default fill value
            keep = tuple(slice(0, min(o, n)) for o, n in zip(self._a.shape, shape))
            a[keep] = self._a[keep]
            self._a = a

        def __setitem__(self, key, value):
            self._a[key] = value

        def __getitem__(self, key):
            return self._a[key]

    class Group(dict):
        def create_dataset(self, name, shape, maxshape, chunks, dtype):
            if name in self:
                raise ValueError("dataset %s exists" % name)
            ds = MemoryStore.Dataset(shape, maxshape, chunks, dtype)
            self[name] = ds
            return ds

    def __init__(self, path):
        self.path = path
        self.groups = {}

    def create_group(self, name):
        g = MemoryStore.Group()
        self.groups[name.strip("/")] = g
        return g

    def __getitem__(self, name):
        parts = name.strip("/").split("/")
        g = self.groups[parts[0]]
        return g if len(parts) == 1 else g[parts[1]]

    def close(self):
        if self.path:
            np.savez(self.path, **{"%s/%s" % (g, n): ds[...] for g, grp in self.groups.items()
                                   for n, ds in grp.items()})


def open_file(path):
    """h5py.File(path, "w") when h5py is importable, else a MemoryStore that
    saves to path with the suffix .npz."""
    try:
        import h5py
    except ImportError:
        return MemoryStore(os.path.splitext(path)[0] + ".npz")
    return h5py.File(path, "w")


class ReferenceOutput:
    """The reference's output file for batches of n_threads sites.

    spec: the dyad's DyadSpec; outputs: the "Outputs" names; n_iter:
    time_max / dt; iterations: {"Dyad": it, "Cytosol": it, "SR": it,
    "Channels": it, "IonsChannels": it} ("Iterations for ..."); n_scenarios:
    sites over all batches.  pools: the config.PoolSpec attached to the dyad
    (its cytosol pool names the /Cytosol datasets), box: the
    config.CytosolBoxSpec attached, if any."""

    def __init__(self, file, spec, outputs, n_iter, iterations, n_scenarios, pools=None, box=None):
        self.f = file
        self.spec = spec
        self.outputs = list(outputs)
        self.S = spec.n_sites
        self.M = spec.max_channels()
        self.n_scenarios = int(n_scenarios)
        self.it = {g: int(iterations[g]) for g in GROUPS}
        self.idx_max = {g: int(n_iter) // self.it[g] + 1 for g in GROUPS}
        self.pools = pools
        self.box = box
        grid = [spec.nx, spec.ny] + ([spec.nz] if spec.dim == 3 else [])
        self.dims = {"Dyad": grid, "SR": [self.M], "Channels": [self.M], "IonsChannels": [self.M]}
        if box is not None:
            self.dims["Cytosol"] = [box.nz, box.ny, box.nx]
        elif pools is not None and pools.cytosol_pool:
            self.dims["Cytosol"] = []
        self.groups = {g: self.f.create_group("/" + g) for g in GROUPS}
        self.created = {g: False for g in GROUPS}

    # -- what each group holds --------------------------------------------
    def _selected(self, names_kinds):
        """(name, index) pairs of the reference's selection: ions whose name
        is in outputs, bound fields whose buffer's name is."""
        return [(nm, k) for k, (nm, ion, buf) in enumerate(names_kinds)
                if (buf is None and ion in self.outputs) or (buf is not None and buf in self.outputs)]

    def _dyad_data(self, dyad):
        sp = self.spec
        kinds = [(i.name, i.name, None) for i in sp.ions]
        kinds += [(sp.ions[i].name + "-" + sp.buffers[u].name, sp.ions[i].name, sp.buffers[u].name)
                  for i, u, *_r in sp.bound_fields()]
        shape = (self.S,) + tuple(self.dims["Dyad"])
        return {nm: dyad.GetField(k).reshape(shape) for nm, k in self._selected(kinds)}

    def _cytosol_data(self, dyad):
        sp = self.spec
        if self.box is not None:
            names = self.box.field_names(sp)
            ions = {i.name for i in sp.ions}
            kinds = [(nm, nm if nm in ions else nm.split("-")[0], None if nm in ions else nm.split("-", 1)[1])
                     for nm in names]
            return {nm: dyad.GetCytosolBox(k) for nm, k in self._selected(kinds)}
        _, ci, cb = dyad.GetPools()
        kinds = [(i.name, i.name, None) for i in sp.ions]
        bufs = list(sp.buffers) + list(self.pools.cyto_buffers)
        bound = [(ion.name + "-" + b.name, ion.name, b.name) for ion in sp.ions for b in bufs
                 if ion.name in b.kinetics]
        values = [ci[k] for k in range(sp.n_ions)] + [cb[k] for k in range(len(bound))]
        return {nm: values[k] for nm, k in self._selected(kinds + bound)}

    def _sr_data(self, dyad, jsr):
        if jsr is None:
            jsr = dyad.GetPools()[0]  # [ion, site]
            jsr = np.repeat(np.asarray(jsr)[:, :, None], self.M, axis=2)
        jsr = np.asarray(jsr, dtype=np.float64).reshape(self.spec.n_ions, self.S, self.M)
        return {i.name: jsr[k] for k, i in enumerate(self.spec.ions) if i.name in self.outputs}

    def _channels_data(self, dyad):
        return {"RyR": np.asarray(dyad.GetChannelsStates(), dtype=np.int32).reshape(self.S, self.M)}

    def _ions_channels_data(self, dyad):
        m = dyad.GetIonsNearChannels()
        return {i.name: np.asarray(m[k]).reshape(self.S, self.M) for k, i in enumerate(self.spec.ions)}

    # -- the reference's write rules ----------------------------------------
    def _write_part(self, g, data, dtype, t, batch):
        grp = self.groups[g]
        n = self.S
        if not self.created[g]:
            shape = [1, n] + list(self.dims[g])
            maxshape = [self.idx_max[g], self.n_scenarios] + list(self.dims[g])
            for name, v in data.items():
                ds = grp.create_dataset(name, shape=tuple(shape), maxshape=tuple(maxshape), chunks=tuple(shape),
                                        dtype=dtype)
                ds[0:1, 0:n] = np.asarray(v, dtype=dtype).reshape([1] + shape[1:])
            self.created[g] = True
            return
        for name, v in data.items():
            ds = grp[name]
            cur = list(ds.shape)
            if t == cur[0] and cur[0] != self.idx_max[g]:
                cur[0] += 1
                ds.resize(tuple(cur))
            elif batch * n == cur[1]:
                cur[1] += n
                ds.resize(tuple(cur))
            ds[t:t + 1, cur[1] - n:cur[1]] = np.asarray(v, dtype=dtype).reshape([1, n] + list(self.dims[g]))

    def write(self, j, batch, dyad, jsr=None):
        """_WriteData at step j of batch `batch` from a DyadHIP handle (after
        its update).  jsr: the jSR values [ion, site, channel] passed to
        Update when no jSR pool is attached."""
        if j % self.it["Dyad"] == 0:
            self._write_part("Dyad", self._dyad_data(dyad), np.float64, j // self.it["Dyad"], batch)
        if "Cytosol" in self.dims and j % self.it["Cytosol"] == 0:
            self._write_part("Cytosol", self._cytosol_data(dyad), np.float64, j // self.it["Cytosol"], batch)
        if j % self.it["SR"] == 0:
            self._write_part("SR", self._sr_data(dyad, jsr), np.float64, j // self.it["SR"], batch)
        if j % self.it["Channels"] == 0:
            self._write_part("Channels", self._channels_data(dyad), np.int32, j // self.it["Channels"], batch)
        if j % self.it["IonsChannels"] == 0:
            self._write_part("IonsChannels", self._ions_channels_data(dyad), np.float64,
                             j // self.it["IonsChannels"], batch)

    def close(self):
        self.f.close()


def iterations_from_json(j):
    """The "Iterations for ..." strides of a reference parameter file."""
    return {"Dyad": j["Iterations for Dyad"], "Cytosol": j["Iterations for Cytosol"], "SR": j["Iterations for SR"],
            "Channels": j["Iterations for IonChannels"], "IonsChannels": j["Iterations for IonsNearChannels"]}
