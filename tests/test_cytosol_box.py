"""The 3D cytosol box coupled at the dyad's rim.

The geometry has no counterpart in the reference (its coupled solver,
DyadRD2DwCytosol.cu, is a 2D composite grid), so the oracle's box is pinned
by what the mathematics fixes: exact conservation of each ion's total over
dyad + box, a discrete eigenmode of the box stencil decaying by its closed
form factor, exact interpolation of a linear field onto the rim, and the
one-step closed form of the reaction and SERCA terms.  The GPU path is then
compared with the oracle on the same inputs (FP64 1e-9, FP32 1e-4) and, at
the full per-site shape of the coupled workload, through conservation.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from helpers import rel_err, spec  # noqa: E402

from casim_amd import abi, workloads  # noqa: E402
from casim_amd.config import Buffer, CytosolBoxSpec, Ion, PoolSpec  # noqa: E402

FP64_TOL = 1e-9
FP32_TOL = 1e-4


def say(*a):
    print("[box]", *a, flush=True)


def species():
    ions = [Ion("Calcium", 220000.0, 0.1)]
    buffers = [Buffer("A", 5000.0, {"Calcium": (140000.0, 0.225, 45.0)}),
               Buffer("B", 24.0, {"Calcium": (25000.0, 0.1, 0.038)}),
               Buffer("F", 1000.0, {"Calcium": (0.0, 0.115, 1.0)})]
    return ions, buffers


def constant_pool():
    return PoolSpec(jsr_constant=True, jsr_cq_total=0.0, jsr_volume=1.0, jsr_c0=[1000.0], jsr_cq_sites=[0.0],
                    jsr_k_cq=[1.0], jsr_tau_refill=[1.0], nsr_c=[1000.0])


def depleting_pool():
    return PoolSpec(jsr_constant=False, jsr_cq_total=10000.0, jsr_volume=2e5, jsr_c0=[1000.0], jsr_cq_sites=[1.0],
                    jsr_k_cq=[800.0], jsr_tau_refill=[100.0], nsr_c=[1000.0])


def small_case(dim=2, n_sites=2, clamped=True, far=False, serca=None, substeps=5, precision=abi.CASIM_FP64):
    """A 40 nm 2D dyad (21 x 21 at h = 2) or a 3D dyad of 11 x 11 x 5 at h = 4,
    centred on a 9 x 9 x 5 box at h = 10 nm."""
    ions, buffers = species()
    if dim == 2:
        sp = spec(dim=2, nx=21, ny=21, h=2.0, ions=ions, buffers=buffers, n_sites=n_sites, clamped=clamped,
                  channels=[(20.0, 20.0), (16.0, 22.0), (26.0, 14.0)], z=16.0, precision=precision)
    else:
        buffers[2] = Buffer("F", 1000.0, {"Calcium": (0.0, 0.115, 1.0)}, faces_only=True)
        sp = spec(dim=3, nx=11, ny=11, nz=5, h=4.0, ions=ions, buffers=buffers, n_sites=n_sites, clamped=clamped,
                  channels=[(20.0, 20.0), (16.0, 24.0)], precision=precision)
    box = CytosolBoxSpec.centred(sp, 9, 9, 5, 10.0, substeps,
                                 [buffers[0], buffers[1], workloads.troponin()], far_dirichlet=far,
                                 serca=serca or {})
    return sp, box


def oracle_run(oracle_mod, sp, box, pool=None):
    ora = oracle_mod.OracleHybridRun(sp, pool or constant_pool())
    ora.attach_box(box)
    ora.reset()
    return ora


def dyad_volume(sp):
    """Volume of each dyad node that the explicit step conserves: interior
    nodes only (the rim is Dirichlet), dx dy z in 2D, dx dy dz with half
    weight on the two z-planes in 3D; [x, y, z]."""
    w = np.zeros((sp.nx, sp.ny, sp.nz))
    w[1:-1, 1:-1, :] = 1.0
    if sp.dim == 3:
        w[:, :, 0] *= 0.5
        w[:, :, -1] *= 0.5
        return w * sp.dx * sp.dy * sp.dz
    return w * sp.dx * sp.dy * sp.z


def total_calcium(sp, box, dyad_fields, box_fields):
    """Each site's Ca2+ amount (free + bound), dyad interior + box, uM nm^3."""
    vd = dyad_volume(sp)
    vb = box.node_weights() * box.h ** 3
    tot = np.zeros(sp.n_sites)
    for f in dyad_fields:
        tot += (f * vd[None]).reshape(sp.n_sites, -1).sum(1)
    for f in box_fields:
        tot += (f * vb[None]).reshape(sp.n_sites, -1).sum(1)
    return tot


def perturb(gen, sp, box, set_dyad, set_box, get_box):
    """Raise the dyad's free Ca2+ and scatter the box's fields (bound below
    their totals) so that every exchange path carries flux."""
    d0 = np.full((sp.n_sites, sp.nx, sp.ny, sp.nz), 0.1)
    d0[:, 3:-3, 3:-3, :] = 5.0 + gen.random((sp.n_sites, sp.nx - 6, sp.ny - 6, sp.nz))
    set_dyad(0, d0)
    shape = (sp.n_sites, box.nz, box.ny, box.nx)
    for f in range(len(box.field_names(sp))):
        base = get_box(f)
        set_box(f, base * (0.5 + gen.random(shape)))


# ---- oracle pins (CPU) -----------------------------------------------------

@pytest.mark.parametrize("dim", [2, 3])
def test_oracle_box_conserves_calcium(oracle_mod, dim):
    """No channel current, no SERCA, no-flux faces: Ca2+ free + bound over the
    dyad's interior and the box is constant at every box step, while more
    than a part in 10^4 of it moves between the two."""
    sp, box = small_case(dim=dim, far=False)
    ora = oracle_run(oracle_mod, sp, box)
    gen = np.random.default_rng(1)
    perturb(gen, sp, box, ora.dyad.set_field, ora.set_box_field, ora.box_field)
    ora.update()
    nf = sp.n_ions + sp.n_bound
    nbf = len(box.field_names(sp))

    def totals():
        return (total_calcium(sp, box, [ora.dyad.field(f) for f in range(nf)], []),
                total_calcium(sp, box, [], [ora.box_field(f) for f in range(nbf)]))

    d0, b0 = totals()
    for _ in range(4):
        ora.step(2e-6, box.substeps)
        d1, b1 = totals()
        assert np.max(np.abs((d1 + b1) - (d0 + b0)) / (d0 + b0)) < 1e-12
    assert np.min(np.abs(d1 - d0) / (d0 + b0)) > 1e-4


def test_oracle_box_rim_is_bilinear(oracle_mod):
    """A box field linear in x and y reaches the dyad's rim exactly: the rim
    node at dyad (i, j) reads a + b X + c Y with X = origin_x + i dx."""
    sp, box = small_case(dim=2)
    ora = oracle_run(oracle_mod, sp, box)
    X = np.arange(box.nx) * box.h
    Y = np.arange(box.ny) * box.h
    lin = 0.3 + 0.002 * X[None, None, None, :] + 0.0035 * Y[None, None, :, None] + np.zeros((sp.n_sites, box.nz, 1, 1))
    ora.set_box_field(0, lin)
    ora.update()
    ca = ora.dyad.field(0)
    i = np.arange(sp.nx)
    j = np.arange(sp.ny)
    want = 0.3 + 0.002 * (box.origin[0] + i[:, None] * sp.dx) + 0.0035 * (box.origin[1] + j[None, :] * sp.dy)
    rim = np.zeros((sp.nx, sp.ny), bool)
    rim[0, :] = rim[-1, :] = rim[:, 0] = rim[:, -1] = True
    for s in range(sp.n_sites):
        assert np.max(np.abs(ca[s, :, :, 0][rim] - want[rim])) < 1e-14
        assert np.all(ca[s, 1:-1, 1:-1, 0] == ca[s, 1, 1, 0])  # the interior is untouched


def test_oracle_box_eigenmode(oracle_mod):
    """Pure diffusion in the box (the dyad's D = 0, so no rim flux): the mode
    sin(pi x/(nx-1)) sin(pi y/(ny-1)) cos(pi z/(2(nz-1))), zero on the held
    faces and symmetric at the no-flux z = 0 face, is an eigenvector of the
    7-point stencil; each box step multiplies it by 1 - dt D lambda with
    lambda = sum over axes of (2 - 2 cos theta) / h^2."""
    ions = [Ion("Calcium", 0.0, 0.1)]
    sp = spec(dim=2, nx=21, ny=21, h=2.0, ions=ions, buffers=[], n_sites=1, z=16.0)
    D = 220000.0
    box = CytosolBoxSpec.centred(sp, 9, 11, 7, 10.0, 30, [], far_dirichlet=True, ion_D={"Calcium": D})
    ora = oracle_run(oracle_mod, sp, box)
    x, y, z = np.arange(box.nx), np.arange(box.ny), np.arange(box.nz)
    tx, ty, tz = np.pi / (box.nx - 1), np.pi / (box.ny - 1), np.pi / (2 * (box.nz - 1))
    mode = (np.cos(tz * z)[:, None, None] * np.sin(ty * y)[None, :, None] * np.sin(tx * x)[None, None, :])
    eps = 0.05
    ora.set_box_field(0, (0.1 + eps * mode)[None])
    ora.update()
    dt = 2e-6
    dtb = dt * box.substeps
    lam = sum((2 - 2 * np.cos(t)) / box.h ** 2 for t in (tx, ty, tz))
    g = 1 - dtb * D * lam
    for n in range(1, 6):
        ora.step(dt, box.substeps)
        got = (ora.box_field(0)[0] - 0.1) / eps
        assert np.max(np.abs(got - g ** n * mode)) < 1e-11, n
    assert g ** 5 < 0.9  # the mode did decay


def test_oracle_box_reaction_and_serca_one_step(oracle_mod):
    """A uniform box away from the dyad's face: one box step gives
    Ca' = c + dt (R - Jmax (c^2/(c^2+K^2) - c0^2/(c0^2+K^2))) and
    CaB' = B - dt R_B with R_B = koff B - kon c (Btot - B) for each buffer."""
    jmax, K = 0.2, 0.184
    sp, box = small_case(dim=2, far=True, serca={"Calcium": (jmax, K)})
    ora = oracle_run(oracle_mod, sp, box)
    c, bvals = 0.7, [80.0, 3.0, 20.0]
    shape = (sp.n_sites, box.nz, box.ny, box.nx)
    ora.set_box_field(0, np.full(shape, c))
    for k, b in enumerate(bvals):
        ora.set_box_field(1 + k, np.full(shape, b))
    ora.update()
    dt = 2e-6
    dtb = dt * box.substeps
    ora.step(dt, box.substeps)
    R = []
    for (i, u, D, kon, koff, _), b in zip(box.bound_fields(sp), bvals):
        R.append(koff * b - kon * c * (box.buffers[u].total - b))
    c0 = 0.1
    serca = jmax * (c * c / (c * c + K * K) - c0 * c0 / (c0 * c0 + K * K))
    want_c = c + dtb * (sum(R) - serca)
    inner = (slice(None), slice(2, box.nz - 1), slice(1, box.ny - 1), slice(1, box.nx - 1))
    assert np.max(np.abs(ora.box_field(0)[inner] - want_c)) < 1e-15 * 10
    for k, b in enumerate(bvals):
        assert np.max(np.abs(ora.box_field(1 + k)[inner] - (b - dtb * R[k]))) < 1e-12
    # the held faces keep their values
    assert np.all(ora.box_field(0)[:, -1] == c)


def test_oracle_box_refuses_rim_outside(oracle_mod):
    sp, box = small_case(dim=2)
    box.origin = (5.0, 20.0)  # the rim's support reaches the x = 0 face
    ora = oracle_mod.OracleHybridRun(sp, constant_pool())
    with pytest.raises(ValueError):
        ora.attach_box(box)


def test_box_struct_fields():
    """The box's bound fields pair with the dyad's by (ion, buffer name);
    the immobile dyad buffer F has no partner and troponin none in the dyad."""
    sp = workloads.coupled_dyads(2)
    box = workloads.coupled_box(sp)
    assert box.field_names(sp) == ["Calcium", "Calcium-A", "Calcium-B", "Calcium-Tn"]
    assert [r[5] for r in box.bound_fields(sp)] == [0, 1, -1]
    c, keep = box.to_struct(sp)
    assert (c.nx, c.ny, c.nz, c.substeps, c.far_dirichlet, c.n_bound) == (51, 51, 51, 50, 1, 3)
    assert box.origin == (300.0, 300.0)
    w = box.node_weights()
    assert w.shape == (51, 51, 51) and w.sum() == 50.0 ** 3


def test_library_exports_box_symbols():
    lib = abi.load()
    for name in ("casim_dyad_attach_cytosol_box", "casim_dyad_cytosol_box_fields", "casim_dyad_get_cytosol_box",
                 "casim_dyad_set_cytosol_box"):
        assert hasattr(lib, name)


# ---- GPU against the oracle -------------------------------------------------

def gpu_pair(oracle_mod, sp, box, pool, perturbed=False, seed=3):
    from casim_amd.dyad import DyadHIP
    ora = oracle_run(oracle_mod, sp, box, pool)
    dev = DyadHIP(sp)
    dev.AttachPools(pool)
    dev.AttachCytosolBox(box)
    if perturbed:
        perturb(np.random.default_rng(seed), sp, box, ora.dyad.set_field, ora.set_box_field, ora.box_field)
        for f in range(sp.n_ions + sp.n_bound):
            dev.SetField(f, ora.dyad.field(f))
        for f in range(len(box.field_names(sp))):
            dev.SetCytosolBox(f, ora.box_field(f))
    ora.init_opening()
    dev.InitOpeningAll()
    ora.update()
    dev.UpdatePools()
    return ora, dev


def compare_all(ora, dev, sp, box, tol):
    for f in range(sp.n_ions + sp.n_bound):
        assert rel_err(dev.GetField(f), ora.dyad.field(f)) < tol, ("dyad field", f)
    for f in range(len(box.field_names(sp))):
        assert rel_err(dev.GetCytosolBox(f), ora.box_field(f)) < tol, ("box field", f)
    assert np.array_equal(dev.GetChannelsStates(), ora.dyad.channel_states())


@pytest.mark.gpu
@pytest.mark.parametrize("dim,precision,substeps", [(2, abi.CASIM_FP64, 5), (3, abi.CASIM_FP64, 5),
                                                    (2, abi.CASIM_FP32, 5), (2, abi.CASIM_FP64, 4),
                                                    (3, abi.CASIM_FP64, 4)])
def test_gpu_box_matches_oracle(gpu, oracle_mod, dim, precision, substeps):
    """Gating channels (FP64; in FP32 the opened channel stays open, as a
    rounding-level difference can flip a draw), a depleting jSR, SERCA and
    held faces: dyad fields, box fields, channel histories and the jSR after
    several calls of different lengths (graphs of both box-buffer parities).
    Odd substeps run one step per pass, even substeps two."""
    say("box parity", dim, precision, substeps)
    fp64 = precision == abi.CASIM_FP64
    sp, box = small_case(dim=dim, n_sites=3, clamped=not fp64, far=True, serca={"Calcium": (0.2, 0.184)},
                         precision=precision, substeps=substeps)
    if fp64:
        sp.channel_types[0].params["kopen0"] *= 1e4
        sp.channel_types[0].params["kclose0"] *= 10
    ora, dev = gpu_pair(oracle_mod, sp, box, depleting_pool(), perturbed=True)
    tol = FP64_TOL if precision == abi.CASIM_FP64 else FP32_TOL
    dt = 2e-6
    assert dev.StepsPerPass() == (2 if substeps % 2 == 0 else 1)
    for k in (1, 2, 3, 1, 4):
        k *= substeps
        dev.Step(dt, k)
        ora.step(dt, k)
        if precision == abi.CASIM_FP64:
            assert np.array_equal(dev.GetChannelsStates(), ora.dyad.channel_states()), k
    compare_all(ora, dev, sp, box, tol)
    assert rel_err(dev.GetPools()[0], ora.jsr()) < tol
    assert (dev.GetChannelsStates() >= 0).any()


@pytest.mark.gpu
@pytest.mark.parametrize("dim,precision", [(2, abi.CASIM_FP64), (3, abi.CASIM_FP64), (2, abi.CASIM_FP32)])
def test_gpu_box_two_step_passes_equal_single_steps(gpu, dim, precision):
    """With a box of even substeps, pairs of steps per pass (the middle
    step's rim effluxes from the ring values the pass writes) give the
    one-step schedule's results bit for bit: dyad and box fields, channel
    histories, RNG and jSR, with fast gating."""
    from casim_amd.dyad import DyadHIP
    say("box pairs", dim, precision)
    sp, box = small_case(dim=dim, n_sites=3, clamped=False, far=True, serca={"Calcium": (0.2, 0.184)},
                         precision=precision, substeps=4)
    sp.channel_types[0].params["kopen0"] *= 1e5
    sp.channel_types[0].params["kclose0"] *= 100
    devs = []
    for steps in (1, 2):
        d = DyadHIP(sp)
        d.AttachPools(depleting_pool())
        d.AttachCytosolBox(box)
        d.SetStepsPerPass(steps)
        perturb(np.random.default_rng(7), sp, box, d.SetField, d.SetCytosolBox, d.GetCytosolBox)
        d.InitOpeningAll()
        d.UpdatePools()
        for k in (4, 8, 12, 4):
            d.Step(2e-6, k)
        devs.append(d)
    one, two = devs
    assert one.StepsPerPass() == 1 and two.StepsPerPass() == 2
    for f in range(sp.n_ions + sp.n_bound):
        assert np.array_equal(one.GetField(f), two.GetField(f)), f
    for f in range(one.CytosolBoxFields()):
        assert np.array_equal(one.GetCytosolBox(f), two.GetCytosolBox(f)), f
    st = one.GetChannelsStates()
    assert np.array_equal(st, two.GetChannelsStates())
    assert np.array_equal(one.GetRngStates(), two.GetRngStates())
    assert np.array_equal(one.GetPools()[0], two.GetPools()[0])


@pytest.mark.gpu
def test_gpu_box_rejects_bad_calls(gpu):
    from casim_amd.dyad import DyadHIP
    sp, box = small_case(dim=2)
    dev = DyadHIP(sp)
    dev.AttachPools(constant_pool())
    dev.AttachCytosolBox(box)
    assert dev.CytosolBoxFields() == 4
    dev.InitOpeningAll()
    dev.UpdatePools()
    with pytest.raises(abi.CasimError):
        dev.Step(2e-6, 7)  # not a multiple of the substeps
    with pytest.raises(abi.CasimError):
        dev.RunRD(2e-6)
    with pytest.raises(abi.CasimError):
        dev.AttachCytosolBox(box)
    with pytest.raises(abi.CasimError):
        dev.GetCytosolBox(9)
    assert dev.StepsPerPass() == 1
    # a box row must fit one wave (nx <= 64), and the rim's support must lie inside the face
    for bad in (CytosolBoxSpec.centred(sp, 65, 9, 5, 10.0, 5, box.buffers),
                CytosolBoxSpec(9, 9, 5, 10.0, (5.0, 20.0), 5, box.buffers)):
        other = DyadHIP(sp)
        with pytest.raises(abi.CasimError):
            other.AttachCytosolBox(bad)
        other.close()


@pytest.mark.gpu
def test_gpu_box_conserves_calcium_full_shape(gpu):
    """The coupled workload's per-site shape (2D 201 x 201 dyad, 51^3 box,
    50 substeps) on 16 sites with channels closed, SERCA off and no-flux
    faces: Ca2+ over dyad + box is conserved to 1e-11 over two box steps
    while Ca2+ moves between the two."""
    from casim_amd.dyad import DyadHIP
    say("box conservation, full shape")
    sp = workloads.coupled_dyads(16, clamped=True)
    box = workloads.coupled_box(sp)
    box.far_dirichlet = False
    dev = DyadHIP(sp)
    dev.AttachPools(constant_pool())
    dev.AttachCytosolBox(box)
    gen = np.random.default_rng(5)
    perturb(gen, sp, box, dev.SetField, dev.SetCytosolBox, dev.GetCytosolBox)
    dev.UpdatePools()
    nf = sp.n_ions + sp.n_bound
    nbf = dev.CytosolBoxFields()

    def totals():
        return (total_calcium(sp, box, [dev.GetField(f) for f in range(nf)], []),
                total_calcium(sp, box, [], [dev.GetCytosolBox(f) for f in range(nbf)]))

    d0, b0 = totals()
    dev.Step(2e-6, 2 * box.substeps)
    d1, b1 = totals()
    assert np.max(np.abs((d1 + b1) - (d0 + b0)) / (d0 + b0)) < 1e-11
    assert np.min(np.abs(d1 - d0) / (d0 + b0)) > 1e-6


@pytest.mark.gpu
def test_gpu_box_full_shape_matches_oracle(gpu, oracle_mod):
    """The coupled workload's own shapes (2D 201 x 201 dyad, 51^3 box, 50
    substeps, 64 RyRs, depleting jSR) on 2 sites, fast gating and perturbed
    fields so that the rim exchange carries flux from the first step: after
    100 steps (two box steps, pairs of steps per pass) every dyad and box
    field matches the oracle to 1e-9 and the channel histories are equal."""
    from casim_amd.dyad import DyadHIP  # noqa: F401
    say("box full shape vs oracle")
    sp = workloads.coupled_dyads(2)
    sp.channel_types[0].params["kopen0"] *= 1e5
    sp.channel_types[0].params["kclose0"] *= 1e3
    box = workloads.coupled_box(sp)
    ora, dev = gpu_pair(oracle_mod, sp, box, workloads.coupled_pools(), perturbed=True, seed=9)
    assert dev.StepsPerPass() == 2
    box0 = ora.box_field(0).copy()
    st0 = ora.dyad.channel_states().copy()
    for _ in range(2):
        dev.Step(sp.meta["dt_ms"], 50)
        ora.step(sp.meta["dt_ms"], 50)
        assert np.array_equal(dev.GetChannelsStates(), ora.dyad.channel_states())
    compare_all(ora, dev, sp, box, FP64_TOL)
    assert rel_err(dev.GetPools()[0], ora.jsr()) < FP64_TOL
    assert not np.array_equal(ora.box_field(0), box0)  # the box took up the rim's flux
    assert not np.array_equal(ora.dyad.channel_states(), st0)  # channels toggled


@pytest.mark.gpu
def test_gpu_box_profiled_steps_equal_graph_steps(gpu):
    """casim_dyad_step_profiled (eager launches with HIP events, what the
    benchmark times) gives the graph path's results bit for bit with a box,
    in pairs and in single steps."""
    from casim_amd.dyad import DyadHIP
    for substeps in (4, 5):
        sp, box = small_case(dim=2, n_sites=3, clamped=False, far=True, substeps=substeps)
        sp.channel_types[0].params["kopen0"] *= 1e5
        devs = []
        for profiled in (False, True):
            d = DyadHIP(sp)
            d.AttachPools(depleting_pool())
            d.AttachCytosolBox(box)
            d.InitOpeningAll()
            d.UpdatePools()
            for k in (2, 1, 3):
                if profiled:
                    d.StepProfiled(2e-6, k * substeps)
                else:
                    d.Step(2e-6, k * substeps)
            devs.append(d)
        a, b = devs
        for f in range(sp.n_ions + sp.n_bound):
            assert np.array_equal(a.GetField(f), b.GetField(f)), (substeps, f)
        for f in range(a.CytosolBoxFields()):
            assert np.array_equal(a.GetCytosolBox(f), b.GetCytosolBox(f)), (substeps, f)
        assert np.array_equal(a.GetChannelsStates(), b.GetChannelsStates())


@pytest.mark.gpu
@pytest.mark.parametrize("dim", [2, 3])
def test_gpu_box_split_pipelining_equals_one_stream(gpu, dim):
    """Split pipelining with a cytosol box (halves rejoin for every box step)
    gives the one-stream results bit for bit: dyad and box fields, channels,
    RNG, jSR and release statistics."""
    from casim_amd.dyad import DyadHIP
    sp, box = small_case(dim=dim, n_sites=5, clamped=False, far=True, substeps=4)
    sp.channel_types[0].params["kopen0"] *= 1e5
    devs = []
    for split in (0, 1):
        d = DyadHIP(sp)
        d.SetSplit(split)
        d.AttachPools(depleting_pool())
        d.AttachCytosolBox(box)
        perturb(np.random.default_rng(11), sp, box, d.SetField, d.SetCytosolBox, d.GetCytosolBox)
        d.InitOpeningAll()
        d.UpdatePools()
        for k in (4, 8, 12):
            d.Step(2e-6, k)
        assert d.StepsPerPass() == 2
        devs.append(d)
    one, two = devs
    for f in range(sp.n_ions + sp.n_bound):
        assert np.array_equal(one.GetField(f), two.GetField(f)), f
    for f in range(one.CytosolBoxFields()):
        assert np.array_equal(one.GetCytosolBox(f), two.GetCytosolBox(f)), f
    assert np.array_equal(one.GetChannelsStates(), two.GetChannelsStates())
    assert np.array_equal(one.GetRngStates(), two.GetRngStates())
    assert np.array_equal(one.GetPools()[0], two.GetPools()[0])
    assert np.array_equal(one.GetReleaseStats(), two.GetReleaseStats())


@pytest.mark.gpu
@pytest.mark.parametrize("substeps", [None, 4, 5])
def test_gpu_long_calls_equal_short_calls(gpu, substeps):
    """A call of more steps than one captured graph holds (2048, or the most
    whole box steps below it) replays several graphs, the last one shorter;
    its results equal those of short calls bit for bit: dyad and box fields,
    channel and RNG states, jSR and release statistics, with fast gating and
    a depleting jSR, without a box (two-step passes, split halves, an odd
    last step) and with boxes of even and odd substeps."""
    from casim_amd.dyad import DyadHIP
    say("long calls", substeps)
    sp, box = small_case(dim=2, n_sites=3, clamped=False, far=True, substeps=substeps or 4)
    sp.channel_types[0].params["kopen0"] *= 1e4
    sp.channel_types[0].params["kclose0"] *= 10
    long_call, short_calls = (4101, [1000] * 4 + [101]) if substeps is None else (4100, [100] * 41)
    devs = []
    for calls in ([long_call], short_calls):
        d = DyadHIP(sp)
        d.AttachPools(depleting_pool())
        if substeps:
            d.AttachCytosolBox(box)
        d.InitOpeningAll()
        d.UpdatePools(*([] if substeps else sp.rim_arrays()))
        for k in calls:
            d.Step(2e-6, k)
        devs.append(d)
    one, many = devs
    for f in range(sp.n_ions + sp.n_bound):
        assert np.array_equal(one.GetField(f), many.GetField(f)), f
    if substeps:
        for f in range(one.CytosolBoxFields()):
            assert np.array_equal(one.GetCytosolBox(f), many.GetCytosolBox(f)), f
    assert np.array_equal(one.GetChannelsStates(), many.GetChannelsStates())
    assert np.array_equal(one.GetRngStates(), many.GetRngStates())
    assert np.array_equal(one.GetPools()[0], many.GetPools()[0])
    assert np.array_equal(one.GetReleaseStats(), many.GetReleaseStats())
    assert one.GetReleaseStats()[:, 0].max() > 0  # channels opened: the runs are not trivial
