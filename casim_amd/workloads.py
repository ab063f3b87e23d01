"""Synthetic release-site workloads.

single_dyad(): one 3D cleft of 240 x 240 x 16 nm at h = 4 nm (61 x 61 x 5 nodes)
with free Ca2+, a mobile ATP-like buffer A, a mobile calmodulin-like buffer B
and an immobile membrane buffer F on the two z-faces; 25 RyRs on a 5 x 5
lattice of 31 nm pitch, SheepCannell-form two-state gating
(kopen = 0.2/ms ([Ca]/15 uM)^2, kclose = 0.5/ms), unitary flux 0.35 pA per mM
of jSR-mouth gradient, jSR fixed at 1 mM, the central RyR opened at t = 0.

batched_dyads(): thousands of independent sites with the same species and
spacing, RyR counts uniform in 8..100 per site, each cluster a random
4-connected accretion on the 31 nm lattice.  A 10 x 10 lattice (279 nm span)
does not fit the 240 nm cleft, so these sites use a 320 x 320 x 16 nm cleft
(81 x 81 x 5 nodes); spacing, species and parameters are unchanged.
"""
import numpy as np

from . import abi
from . import rng
from .config import Buffer, ChannelType, DyadSpec, Ion

FARADAY = 96485.33212  # C/mol
DT_MS = 1e-5  # 10 ns


def unitary_conductance(pA_per_mM=0.35, node_volume_nm3=64.0):
    """Concentration rate at the mouth node per uM of jSR-mouth gradient,
    (uM/ms)/uM, for a unitary current of pA_per_mM pA per mM of gradient
    carried by Ca2+ (z = 2) into one node volume."""
    mol_per_s_per_mM = pA_per_mM * 1e-12 / (2 * FARADAY)
    litres = node_volume_nm3 * 1e-24
    molar_per_s_per_mM = mol_per_s_per_mM / litres  # M/s per mM
    return molar_per_s_per_mM * 1e3 / 1e3  # (uM/ms) per uM: M/s -> 1e3 uM/ms, per mM -> per 1e3 uM


def _species(faces_only_fixed=True):
    ions = [Ion("Calcium", 220000.0, 0.1, 1000.0)]
    buffers = [
        Buffer("A", 5000.0, {"Calcium": (140000.0, 0.225, 45.0)}),
        Buffer("B", 24.0, {"Calcium": (25000.0, 0.1, 0.038)}),
        Buffer("F", 1000.0, {"Calcium": (0.0, 0.115, 1.0)}, faces_only=faces_only_fixed),
    ]
    return ions, buffers


def ryr_type(clamped=False):
    """SheepCannell form of kopen = 0.2/ms ([Ca]/15uM)^2, kclose = 0.5/ms.
    SheepCannell evaluates kopen0 (Ca 1e-3)^n 1e-3, so kopen0 = 0.2/225 * 1e9.
    clamped=True zeroes both rates: channels keep their InitOpening state
    while the draws still happen."""
    p = {"kopen0": 0.0 if clamped else 0.2 / 225.0 * 1e9, "ncaopen": 2.0, "kopenmax": 1e300,
         "kclose0": 0.0 if clamped else 500.0, "ncaclose": 0.0}
    return ChannelType("RyR", "SheepCannell", p, [unitary_conductance()])


def single_dyad(n_sites=32, seed=rng.DEFAULT_SEED, clamped=False, precision=abi.CASIM_FP64,
                faces_only_fixed=True):
    ions, buffers = _species(faces_only_fixed)
    h = 4.0
    lattice = [(120.0 + 31.0 * i, 120.0 + 31.0 * j, 0) for i in range(-2, 3) for j in range(-2, 3)]
    spec = DyadSpec(dim=3, nx=61, ny=61, nz=5, dx=h, dy=h, dz=h, z=16.0, radius=1000.0, voltage=0.0,
                    ions=ions, buffers=buffers, sl={}, start_equilibrium=True,
                    channel_types=[ryr_type(clamped)], sites=[list(lattice) for _ in range(n_sites)],
                    rng_states=rng.stream_states(seed, n_sites), precision=precision,
                    init_channel=[12] * n_sites,
                    meta={"workload": "single-dyad 61x61x5 FP%d, 25 RyRs" % (64 if precision == 0 else 32),
                          "jsr_uM": 1000.0, "dt_ms": DT_MS})
    return spec


def random_cluster(gen, n, side=10):
    """n cells of a side x side lattice grown by 4-connected random accretion
    from the lattice centre; returns lattice (i, j) in growth order."""
    c = side // 2
    cells = [(c, c)]
    taken = {(c, c)}
    frontier = []

    def push(p):
        for d in ((1, 0), (-1, 0), (0, 1), (0, -1)):
            q = (p[0] + d[0], p[1] + d[1])
            if 0 <= q[0] < side and 0 <= q[1] < side and q not in taken and q not in frontier:
                frontier.append(q)

    push((c, c))
    while len(cells) < n:
        q = frontier.pop(int(gen.integers(len(frontier))))
        cells.append(q)
        taken.add(q)
        push(q)
    return cells


def batched_dyads(n_sites=4096, seed=rng.DEFAULT_SEED, geometry_seed=2024, n_min=8, n_max=100,
                  clamped=False, precision=abi.CASIM_FP64, faces_only_fixed=True, extent_nm=320.0):
    ions, buffers = _species(faces_only_fixed)
    h = 4.0
    n = int(round(extent_nm / h)) + 1
    gen = np.random.default_rng(geometry_seed)
    side = 10
    centre = extent_nm / 2.0
    sites = []
    for s in range(n_sites):
        k = int(gen.integers(n_min, n_max + 1))
        cells = random_cluster(gen, k, side)
        sites.append([(centre + 31.0 * (i - (side - 1) / 2.0), centre + 31.0 * (j - (side - 1) / 2.0), 0)
                      for i, j in cells])
    spec = DyadSpec(dim=3, nx=n, ny=n, nz=5, dx=h, dy=h, dz=h, z=16.0, radius=1000.0, voltage=0.0,
                    ions=ions, buffers=buffers, sl={}, start_equilibrium=True,
                    channel_types=[ryr_type(clamped)], sites=sites,
                    rng_states=rng.stream_states(seed, n_sites), precision=precision,
                    init_channel=[0] * n_sites,
                    meta={"workload": "batched %d dyads %dx%dx5 FP%d, RyRs %d..%d"
                          % (n_sites, n, n, 64 if precision == 0 else 32, n_min, n_max),
                          "jsr_uM": 1000.0, "dt_ms": DT_MS})
    return spec


def troponin():
    """Immobile troponin-like cytosol buffer, 70 uM (TnC low-affinity site rates)."""
    return Buffer("Tn", 70.0, {"Calcium": (0.0, 0.0327, 0.0196)})


def coupled_dyads(n_sites=512, seed=rng.DEFAULT_SEED, n_ryr=64, extent_nm=400.0, h=2.0, clamped=False,
                  precision=abi.CASIM_FP64):
    """2D clefts of 400 x 400 nm at h = 2 nm (201 x 201 nodes, element
    thickness 16 nm, so a node holds 64 nm^3 as in the 3D cleft), the species
    of single_dyad with the fixed buffer F on every node, 64 RyRs on an 8 x 8
    lattice of 31 nm pitch, the channel nearest the centre opened at t = 0,
    dt 2 ns (diffusion number 0.44 at h = 2 nm).  Pair with coupled_box() and
    coupled_pools()."""
    ions, buffers = _species(faces_only_fixed=False)
    n = int(round(extent_nm / h)) + 1
    side = int(round(np.sqrt(n_ryr)))
    c = extent_nm / 2.0
    lattice = [(c + 31.0 * (i - (side - 1) / 2.0), c + 31.0 * (j - (side - 1) / 2.0), 0)
               for i in range(side) for j in range(side)]
    centre = min(range(len(lattice)), key=lambda k: (lattice[k][0] - c) ** 2 + (lattice[k][1] - c) ** 2)
    spec = DyadSpec(dim=2, nx=n, ny=n, nz=1, dx=h, dy=h, dz=h, z=16.0, radius=1000.0, voltage=0.0,
                    ions=ions, buffers=buffers, sl={}, start_equilibrium=True,
                    channel_types=[ryr_type(clamped)], sites=[list(lattice) for _ in range(n_sites)],
                    rng_states=rng.stream_states(seed, n_sites), precision=precision,
                    init_channel=[centre] * n_sites,
                    meta={"workload": "coupled %d dyads 2D %dx%d + cytosol box, %d RyRs"
                          % (n_sites, n, n, len(lattice)),
                          "jsr_uM": 1000.0, "dt_ms": 2e-6})
    return spec


def coupled_box(spec, n=51, h=20.0, substeps=50):
    """The 1 x 1 x 1 um cytosol box at h = 20 nm (51^3 nodes) around each
    coupled dyad: Ca2+, the dyad's mobile buffers A and B, immobile
    troponin 70 uM; outer faces held at rest, the dyad centred on the z = 0
    face; one box step per 50 dyad steps (100 ns, diffusion number 0.33)."""
    from .config import CytosolBoxSpec
    mobile = [b for b in spec.buffers if any(k[0] > 0 for k in b.kinetics.values())]
    return CytosolBoxSpec.centred(spec, n, n, n, h, substeps, mobile + [troponin()], far_dirichlet=True)


def coupled_pools():
    """Depleting jSR: 1 mM start, 10 mM calsequestrin-like buffer with Kd
    0.8 mM, refill tau 100 ms from a 1 mM NSR, V_jSR 1e7 nm^3."""
    from .config import PoolSpec
    return PoolSpec(jsr_constant=False, jsr_cq_total=10000.0, jsr_volume=1e7, jsr_c0=[1000.0],
                    jsr_cq_sites=[1.0], jsr_k_cq=[800.0], jsr_tau_refill=[100.0], nsr_c=[1000.0])
