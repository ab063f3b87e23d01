"""Site configurations: CaSim JSON -> flat arrays -> CasimDyadConfig.

Parsing follows what the DyadRDCPU / DyadRD2DCPU constructors read
(DyadRDCPU.cpp:8-161) and DyadChannels::_Initialize (DyadChannels.cpp:17-61).
nlohmann::json iterates object keys in sorted order, which fixes the order of
ions, buffers, channel types and conductances; bound fields are ion-major
over (ion, buffer) pairs (DyadRDCPU.cpp:84-99).
"""
import ctypes as C
import json
import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import abi
from . import rng

MODEL_IDS = {
    "SheepCannell": abi.CASIM_MODEL_SHEEP_CANNELL,
    "TwoStateMWC": abi.CASIM_MODEL_TWO_STATE_MWC,
    "TwoStateaHTG": abi.CASIM_MODEL_TWO_STATE_AHTG,
}
MODEL_PARAMS = {
    "SheepCannell": ["kopen0", "ncaopen", "kopenmax", "kclose0", "ncaclose"],
    "TwoStateMWC": ["fCa", "KCa", "fMg", "KMg", "KO0", "KMgI", "kclose"],
    "TwoStateaHTG": ["fCa", "KCa", "fMg", "KMg", "KO0", "KMgI", "KCL", "KI", "KOL", "kclose"],
}


@dataclass
class Ion:
    name: str
    D: float
    c0: float
    extracell: float = 0.0


@dataclass
class Buffer:
    name: str
    total: float
    kinetics: Dict[str, Tuple[float, float, float]]  # ion -> (D, kon, koff)
    faces_only: bool = False


@dataclass
class ChannelType:
    name: str
    model: str
    params: Dict[str, float]
    conductance: List[float]  # positional, sorted "Conductance" keys


@dataclass
class DyadSpec:
    """Everything casim_dyad_create / oracle_create need, as Python data."""
    dim: int
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    z: float
    radius: float
    voltage: float
    ions: List[Ion]
    buffers: List[Buffer]
    sl: Dict[str, Tuple[float, float, float, float]]
    start_equilibrium: bool
    channel_types: List[ChannelType]
    sites: List[List[Tuple[float, float, int]]]  # per site: (x nm, y nm, type)
    rng_states: np.ndarray  # [S, 6] uint64
    precision: int = abi.CASIM_FP64
    init_channel: Optional[List[int]] = None  # channel opened by InitOpening per site
    meta: Dict = field(default_factory=dict)

    # ---- derived orderings ----
    @property
    def n_sites(self):
        return len(self.sites)

    def bound_fields(self):
        """(ion index, buffer index, D, kon, koff) ion-major (DyadRDCPU.cpp:84-99)."""
        out = []
        for i, ion in enumerate(self.ions):
            for u, b in enumerate(self.buffers):
                if ion.name in b.kinetics:
                    D, kon, koff = b.kinetics[ion.name]
                    out.append((i, u, D, kon, koff))
        return out

    def field_names(self):
        names = [i.name for i in self.ions]
        for i, u, *_ in self.bound_fields():
            names.append(self.ions[i].name + "-" + self.buffers[u].name)
        return names

    @property
    def n_ions(self):
        return len(self.ions)

    @property
    def n_bound(self):
        return len(self.bound_fields())

    @property
    def n_buffers(self):
        return len(self.buffers)

    def ion_index(self, name):
        for k, i in enumerate(self.ions):
            if i.name == name:
                return k
        return -1

    def max_channels(self):
        return max(len(s) for s in self.sites)

    def nodes_per_site(self):
        return self.nx * self.ny * self.nz

    def to_struct(self):
        """(CasimDyadConfig, keepalive) -- keep the second alive while the first is used."""
        S = self.n_sites
        bf = self.bound_fields()
        keep = {}

        def arr(name, values, dtype):
            a = np.ascontiguousarray(np.array(values, dtype=dtype).reshape(-1))
            if a.size == 0:
                a = np.zeros(1, dtype=dtype)
            keep[name] = a
            return a

        cfg = abi.CasimDyadConfig()
        cfg.dim, cfg.nx, cfg.ny, cfg.nz = self.dim, self.nx, self.ny, self.nz
        cfg.dx, cfg.dy, cfg.dz, cfg.z = self.dx, self.dy, self.dz, self.z
        cfg.radius, cfg.voltage = self.radius, self.voltage
        cfg.n_sites = S
        cfg.n_ions, cfg.n_buffers, cfg.n_bound = self.n_ions, self.n_buffers, len(bf)
        cfg.ion_D = abi.dptr(arr("ion_D", [i.D for i in self.ions], np.float64))
        cfg.ion_c0 = abi.dptr(arr("ion_c0", [i.c0 for i in self.ions], np.float64))
        cfg.buffer_total = abi.dptr(arr("buf_total", [b.total for b in self.buffers], np.float64))
        cfg.buffer_faces_only = abi.iptr(arr("faces", [int(b.faces_only) for b in self.buffers], np.int32))
        cfg.bound_ion = abi.iptr(arr("b_ion", [t[0] for t in bf], np.int32))
        cfg.bound_buffer = abi.iptr(arr("b_buf", [t[1] for t in bf], np.int32))
        cfg.bound_D = abi.dptr(arr("b_D", [t[2] for t in bf], np.float64))
        cfg.bound_kon = abi.dptr(arr("b_kon", [t[3] for t in bf], np.float64))
        cfg.bound_koff = abi.dptr(arr("b_koff", [t[4] for t in bf], np.float64))
        cfg.start_equilibrium = int(self.start_equilibrium)
        sl = [self.sl.get(i.name) for i in self.ions]
        cfg.sl_binds = abi.iptr(arr("sl_b", [int(s is not None) for s in sl], np.int32))
        for k, nm in enumerate(["N1", "K1", "N2", "K2"]):
            setattr(cfg, "sl_" + nm, abi.dptr(arr("sl_" + nm, [s[k] if s else 0.0 for s in sl], np.float64)))
        cfg.calcium_index = self.ion_index("Calcium")
        cfg.magnesium_index = self.ion_index("Magnesium")
        T = len(self.channel_types)
        cfg.n_channel_types = T
        params = np.zeros((T, abi.CASIM_MAX_MODEL_PARAMS))
        cond = np.zeros((T, self.n_ions))
        for t, ct in enumerate(self.channel_types):
            for k, pn in enumerate(MODEL_PARAMS[ct.model]):
                params[t, k] = ct.params.get(pn, 0.0)
            for k in range(self.n_ions):
                cond[t, k] = ct.conductance[k] if k < len(ct.conductance) else 0.0
        cfg.type_model = abi.iptr(arr("t_model", [MODEL_IDS[c.model] for c in self.channel_types], np.int32))
        cfg.type_params = abi.dptr(arr("t_params", params, np.float64))
        cfg.type_conductance = abi.dptr(arr("t_g", cond, np.float64))
        off = np.zeros(S + 1, dtype=np.int32)
        off[1:] = np.cumsum([len(s) for s in self.sites])
        cfg.site_channel_offset = abi.iptr(arr("off", off, np.int32))
        flat = [c for s in self.sites for c in s]
        cfg.channel_x = abi.dptr(arr("cx", [c[0] for c in flat], np.float64))
        cfg.channel_y = abi.dptr(arr("cy", [c[1] for c in flat], np.float64))
        cfg.channel_type = abi.iptr(arr("ct", [c[2] for c in flat], np.int32))
        rs = np.ascontiguousarray(self.rng_states, dtype=np.uint64).reshape(S, 6)
        keep["rng"] = rs
        cfg.rng_state = abi.uptr(rs)
        cfg.precision = self.precision
        return cfg, keep

    # ---- reference-layout helpers for Update arguments ----
    def bound_init(self):
        """Equilibrium bound concentrations (DyadRDCPU.cpp:53-81), computed in numpy."""
        out = np.zeros(self.n_bound)
        bf = self.bound_fields()
        for u, b in enumerate(self.buffers):
            rows = [(k, t) for k, t in enumerate(bf) if t[1] == u]
            if not rows or not self.start_equilibrium:
                continue
            m = len(rows)
            A = np.zeros((m, m))
            rhs = np.zeros(m)
            for r, (k, t) in enumerate(rows):
                c = self.ions[t[0]].c0
                A[r, :] = t[3] * c
                A[r, r] += t[4]
                rhs[r] = t[3] * c * b.total
            sol = np.linalg.solve(A, rhs)
            for r, (k, t) in enumerate(rows):
                out[k] = sol[r]
        return out

    def rim_arrays(self):
        """Constant Dirichlet rim at the resting/equilibrium values:
        ions [site + ion*S], bound [site + bound*S]."""
        S = self.n_sites
        ions = np.repeat(np.array([i.c0 for i in self.ions]), S)
        bound = np.repeat(self.bound_init(), S)
        return ions, bound

    def jsr_array(self, values):
        """jSR concentration per ion -> [ch + n_ch*(site + ion*S)]."""
        S, M = self.n_sites, self.max_channels()
        out = np.zeros(M * S * self.n_ions)
        for k, v in enumerate(values):
            out[k * M * S:(k + 1) * M * S] = v
        return out


def element_volume(spec):
    """Node volume the channel currents refer to (DyadRDCPU::GetElementVolume,
    DyadRDCPU.cpp:195-197, 0.5 dx dy dz in 3D; DyadRD2DCPU.cpp:189-191,
    dx dy z in 2D), nm^3."""
    if spec.dim == 3:
        return 0.5 * spec.dx * spec.dy * spec.dz
    return spec.dx * spec.dy * spec.z


def _sorted_items(obj):
    return sorted(obj.items(), key=lambda kv: kv[0])


def from_json(j_dyad, n_sites=1, base_dir=".", seed=rng.DEFAULT_SEED, rng_states=None,
              precision=abi.CASIM_FP64):
    """Build a DyadSpec from the reference's "Dyad" JSON object.

    Every site gets the same channel list, as in the reference.  Site k's RNG
    stream is the k-th stream after SetPackageSeed(seed) unless rng_states is given.
    """
    dim = 2 if j_dyad["Type"] in ("2D", "CPU2D") else 3
    dx, dy, x, y, zz = (float(j_dyad[k]) for k in ("dx", "dy", "x", "y", "z"))
    dz = float(j_dyad.get("dz", 1.0))
    nx, ny = int(x / dx + 1), int(y / dy + 1)  # DyadRDCPU.cpp:18-20 (int truncation)
    nz = int(zz / dz + 1) if dim == 3 else 1
    ions = [Ion(k, float(v["D"]), float(v["Concentration"]), float(v.get("Extracell", 0.0)))
            for k, v in _sorted_items(j_dyad["Ions"])]
    buffers = []
    for k, v in _sorted_items(j_dyad["Buffers"]):
        kin = {ik: (float(iv["D"]), float(iv["kon"]), float(iv["koff"])) for ik, iv in _sorted_items(v["Ions"])}
        for ik in kin:
            if ik not in [i.name for i in ions]:
                raise ValueError("buffer %s binds %s, which is not a dyad ion" % (k, ik))
        buffers.append(Buffer(k, float(v["Total Concentration"]), kin, bool(v.get("Faces only", False))))
    sl = {k: (float(v["N1"]), float(v["K1"]), float(v["N2"]), float(v["K2"]))
          for k, v in _sorted_items(j_dyad.get("Sarcolemma", {}))}
    types = []
    coords = []
    for t, (k, v) in enumerate(_sorted_items(j_dyad["Channels"])):
        model = v["Model"]
        if model not in MODEL_IDS:
            raise ValueError("channel model %s is not supported by the dyad library" % model)
        g = [float(gv) for _, gv in _sorted_items(v["Conductance"])]
        types.append(ChannelType(k, model, {pk: float(pv) for pk, pv in v["Parameters"].items()}, g))
        path = v["Coordinates"]
        if not os.path.isabs(path):
            path = os.path.join(base_dir, path)
        xy = np.loadtxt(path).reshape(-1, 2)
        coords.extend((float(a), float(b), t) for a, b in xy)
    if rng_states is None:
        seed_json = j_dyad.get("Seed")
        if seed_json is not None:
            seed = list(seed_json) + [0] * (6 - len(seed_json))
        rng_states = rng.stream_states(seed, n_sites)
    spec = DyadSpec(dim=dim, nx=nx, ny=ny, nz=nz, dx=dx, dy=dy, dz=dz, z=zz,
                    radius=float(j_dyad["Radius"]), voltage=float(j_dyad.get("Voltage", 0.0)),
                    ions=ions, buffers=buffers, sl=sl,
                    start_equilibrium=j_dyad.get("Start") == "Equilibrium",
                    channel_types=types, sites=[list(coords) for _ in range(n_sites)],
                    rng_states=np.asarray(rng_states, dtype=np.uint64), precision=precision)
    return spec


def load_json(path, n_sites=1, **kw):
    with open(path) as f:
        j = json.load(f)
    return from_json(j["Dyad"], n_sites=n_sites, base_dir=os.path.dirname(os.path.abspath(path)), **kw), j


@dataclass
class PoolSpec:
    """jSR pool, network SR and cytosol around the dyad (SRPool.cpp:4-53,
    NSRConstPool.cpp:3-14, CytosolPool.cpp:3-157), over the dyad's ions."""
    jsr_constant: bool
    jsr_cq_total: float
    jsr_volume: float
    jsr_c0: List[float]
    jsr_cq_sites: List[float]
    jsr_k_cq: List[float]
    jsr_tau_refill: List[float]
    nsr_c: List[float]
    cytosol_pool: bool = False
    cyto_leak_mult: float = 0.0
    cyto_buffers: List[Buffer] = field(default_factory=list)
    serca: Dict[str, Tuple[float, float]] = field(default_factory=dict)  # ion -> (Jmax, Kup)

    def to_struct(self, ions):
        ni = len(ions)
        names = [i.name for i in ions]
        keep = {}

        def arr(name, values, dtype):
            a = np.ascontiguousarray(np.array(values, dtype=dtype).reshape(-1))
            if a.size == 0:
                a = np.zeros(1, dtype=dtype)
            keep[name] = a
            return a

        c = abi.CasimPoolConfig()
        c.jsr_constant = int(self.jsr_constant)
        c.jsr_cq_total = self.jsr_cq_total
        c.jsr_volume = self.jsr_volume
        for nm in ("jsr_c0", "jsr_cq_sites", "jsr_k_cq", "jsr_tau_refill", "nsr_c"):
            setattr(c, nm, abi.dptr(arr(nm, getattr(self, nm), np.float64)))
        c.cytosol_pool = int(self.cytosol_pool)
        c.cyto_leak_mult = self.cyto_leak_mult
        nb = len(self.cyto_buffers)
        c.n_cyto_buffers = nb
        c.cyto_buffer_total = abi.dptr(arr("cbt", [b.total for b in self.cyto_buffers], np.float64))
        D = np.zeros((nb, ni))
        kon = np.zeros((nb, ni))
        koff = np.zeros((nb, ni))
        binds = np.zeros((nb, ni), dtype=np.int32)
        for u, b in enumerate(self.cyto_buffers):
            for ion, (d_, kn, kf) in b.kinetics.items():
                if ion in names:
                    k = names.index(ion)
                    D[u, k], kon[u, k], koff[u, k], binds[u, k] = d_, kn, kf, 1
        c.cyto_buffer_D = abi.dptr(arr("cbD", D, np.float64))
        c.cyto_buffer_kon = abi.dptr(arr("cbkon", kon, np.float64))
        c.cyto_buffer_koff = abi.dptr(arr("cbkoff", koff, np.float64))
        c.cyto_buffer_binds = abi.iptr(arr("cbb", binds, np.int32))
        c.serca_binds = abi.iptr(arr("sb", [int(n in self.serca) for n in names], np.int32))
        c.serca_jmax = abi.dptr(arr("sj", [self.serca.get(n, (0.0, 0.0))[0] for n in names], np.float64))
        c.serca_kup = abi.dptr(arr("sk", [self.serca.get(n, (0.0, 0.0))[1] for n in names], np.float64))
        return c, keep


def pools_from_json(j, spec):
    """PoolSpec from the reference's "jSR", "nSR" and "Cytosol" objects."""
    names = [i.name for i in spec.ions]
    jj = j["jSR"]
    ji = jj.get("Ions", {})
    get = lambda ion, key, dflt: float(ji[ion][key]) if ion in ji else dflt
    nsr = j.get("nSR", {}).get("Ions", {})
    cy = j.get("Cytosol", {})
    if cy.get("Ions"):
        raise ValueError("cytosol-only ions are not supported (CytosolPool indexes them as dyad effluxes)")
    bufs = []
    for k, v in _sorted_items(cy.get("Buffers", {})):
        kin = {ik: (float(iv["D"]), float(iv["kon"]), float(iv["koff"])) for ik, iv in _sorted_items(v["Ions"])}
        bufs.append(Buffer(k, float(v["Total Concentration"]), kin))
    serca = {k: (float(v["Jmax"]), float(v["Kup"])) for k, v in _sorted_items(cy.get("SERCA", {}))}
    return PoolSpec(
        jsr_constant=bool(jj.get("Constant", False)), jsr_cq_total=float(jj.get("Total CQ (uM)", 0.0)),
        jsr_volume=float(jj.get("V jSR (nm3)", 1.0)),
        jsr_c0=[get(n, "Concentration", 0.0) for n in names],
        jsr_cq_sites=[get(n, "Binding sites of CQ", 0.0) for n in names],
        jsr_k_cq=[get(n, "K CQ (uM)", 0.0) for n in names],
        jsr_tau_refill=[get(n, "T refill (ms)", 1.0) for n in names],
        nsr_c=[float(nsr[n]["Concentration (uM)"]) if n in nsr else 0.0 for n in names],
        cytosol_pool=cy.get("Type") == "Pool", cyto_leak_mult=float(cy.get("Leak mult", 0.0)),
        cyto_buffers=bufs, serca=serca)


@dataclass
class CytosolBoxSpec:
    """3D cytosol box around every site, coupled at the dyad's rim
    (CasimCytosolBoxConfig).  Box buffers named like a dyad buffer exchange
    with the dyad's bound field of the same ion at the rim; the dyad's mobile
    bound fields all need such a partner.  origin: where dyad node (0, 0)
    sits on the box's z = 0 face, nm."""
    nx: int
    ny: int
    nz: int
    h: float
    origin: Tuple[float, float]
    substeps: int
    buffers: List[Buffer]
    far_dirichlet: bool = True
    ion_D: Optional[Dict[str, float]] = None  # default: the dyad's
    serca: Dict[str, Tuple[float, float]] = field(default_factory=dict)  # ion -> (Jmax uM/ms, K uM)

    @staticmethod
    def centred(spec, nx, ny, nz, h, substeps, buffers, **kw):
        """A box with the dyad centred on its z = 0 face."""
        ox = ((nx - 1) * h - (spec.nx - 1) * spec.dx) / 2.0
        oy = ((ny - 1) * h - (spec.ny - 1) * spec.dy) / 2.0
        return CytosolBoxSpec(nx, ny, nz, h, (ox, oy), substeps, buffers, **kw)

    def bound_fields(self, spec):
        """(ion index, box buffer index, D, kon, koff, dyad bound field or -1), ion-major."""
        dyad = {(spec.ions[i].name, spec.buffers[u].name): k for k, (i, u, *_r) in enumerate(spec.bound_fields())}
        out = []
        for i, ion in enumerate(spec.ions):
            for u, b in enumerate(self.buffers):
                if ion.name in b.kinetics:
                    D, kon, koff = b.kinetics[ion.name]
                    out.append((i, u, D, kon, koff, dyad.get((ion.name, b.name), -1)))
        return out

    def field_names(self, spec):
        return [i.name for i in spec.ions] + [spec.ions[i].name + "-" + self.buffers[u].name
                                              for i, u, *_r in self.bound_fields(spec)]

    def nodes(self):
        return self.nx * self.ny * self.nz

    def node_weights(self):
        """Finite-volume weight of each node, [z, y, x] in units of h^3: 1/2 per
        no-flux face the node lies on (what the box's mirror stencil conserves)."""
        def axis(n):
            w = np.ones(n)
            w[0] = w[-1] = 0.5
            return w
        return axis(self.nz)[:, None, None] * axis(self.ny)[None, :, None] * axis(self.nx)[None, None, :]

    def to_struct(self, spec):
        keep = {}

        def arr(name, values, dtype):
            a = np.ascontiguousarray(np.array(values, dtype=dtype).reshape(-1))
            if a.size == 0:
                a = np.zeros(1, dtype=dtype)
            keep[name] = a
            return a

        bf = self.bound_fields(spec)
        names = [i.name for i in spec.ions]
        D = [(self.ion_D or {}).get(n, spec.ions[k].D) for k, n in enumerate(names)]
        c = abi.CasimCytosolBoxConfig()
        c.nx, c.ny, c.nz = self.nx, self.ny, self.nz
        c.h = self.h
        c.origin_x, c.origin_y = float(self.origin[0]), float(self.origin[1])
        c.substeps = int(self.substeps)
        c.far_dirichlet = int(self.far_dirichlet)
        c.ion_D = abi.dptr(arr("iD", D, np.float64))
        c.n_buffers = len(self.buffers)
        c.buffer_total = abi.dptr(arr("bt", [b.total for b in self.buffers], np.float64))
        c.n_bound = len(bf)
        c.bound_ion = abi.iptr(arr("bi", [r[0] for r in bf], np.int32))
        c.bound_buffer = abi.iptr(arr("bb", [r[1] for r in bf], np.int32))
        c.bound_D = abi.dptr(arr("bD", [r[2] for r in bf], np.float64))
        c.bound_kon = abi.dptr(arr("bkon", [r[3] for r in bf], np.float64))
        c.bound_koff = abi.dptr(arr("bkoff", [r[4] for r in bf], np.float64))
        c.bound_dyad = abi.iptr(arr("bd", [r[5] for r in bf], np.int32))
        c.serca_jmax = abi.dptr(arr("sj", [self.serca.get(n, (0.0, 0.0))[0] for n in names], np.float64))
        c.serca_k = abi.dptr(arr("sk", [self.serca.get(n, (0.0, 0.0))[1] for n in names], np.float64))
        return c, keep
