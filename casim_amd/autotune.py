"""Autotuner of the state sweep's launch shape.

For one batch shape it times every candidate (kernel form, threads per
workgroup, x-planes per workgroup) by the wall time of the fused step the
solver runs (the captured graph, with split pipelining when it is on) and
keeps the fastest; the sweep launch time from the profiled step entry point,
which brackets each launch with HIP events, is recorded beside it.  The two
can rank candidates differently: two half-batches on two streams overlap,
so the shape with the shortest whole-batch launch is not always the one
with the shortest step.  Results are cached in tuning.json next to this
file, keyed by grid, species layout, precision, site count and steps per
pass; results of the step do not depend on the choice.
"""
import json
import os
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CACHE = os.path.join(HERE, "tuning.json")


def key(spec, steps_per_pass=2):
    return "dim%d_%dx%dx%d_i%d_b%d_f%d_p%d_S%d_k%d" % (
        spec.dim, spec.nx, spec.ny, spec.nz, spec.n_ions, spec.n_bound,
        sum(int(b.faces_only) for b in spec.buffers), spec.precision, spec.n_sites, steps_per_pass)


def candidates(spec):
    P = spec.ny * spec.nz
    inner = spec.nx - 2
    out = [(-1, 256, 0)]
    for th in (64, 128, 192, 256, 320, 384, 448, 512):
        slots = (P + th - 1) // th
        if slots > 4 or th > max(64, ((P + 63) // 64) * 64):
            continue
        if slots > 1 and th * slots - P >= th // 2:
            continue  # more than half a block idle on the last slot
        chunks = {0, (inner + 1) // 2, (inner + 2) // 3, (inner + 3) // 4}
        if spec.n_sites < 2048:  # few sites: shorter chunks give the 256 CUs more workgroups
            chunks |= {(inner + 7) // 8, (inner + 15) // 16}
        for chunk in sorted(chunks):
            for pf in ((1, 2, 3) if slots <= 2 else (1,)):
                out.append((chunk, th, pf))
    return out


def load_cache():
    if os.path.exists(CACHE):
        try:
            with open(CACHE) as f:
                return json.load(f)
        except ValueError:
            return {}
    return {}


def lookup(spec, steps_per_pass=2):
    return load_cache().get(key(spec, steps_per_pass))


def apply(dev, spec):
    """Apply the cached tuning of this shape to a DyadHIP handle, if any."""
    t = lookup(spec, dev.StepsPerPass())
    if t:
        dev.SetTuning(t["planes"], t["threads"], t.get("prefetch", 1))
    return t


def tune(dev, spec, dt, steps=20, warmup=4, save=True, log=None):
    """Time every candidate on `dev` (a DyadHIP after its first Update):
    `steps` profiled steps for the sweep's launch time, then the fastest of
    three calls of 2 `steps` fused steps, after one that captures their
    graph, for the step's wall time, which picks the shape.  The prefetch depth belongs to the
    one-step sweep only, so with two steps per pass each (planes, threads)
    shape is timed once."""
    res = []
    cands = candidates(spec)
    if dev.StepsPerPass() == 2:
        seen = set()
        cands = [c for c in cands if (c[0], c[1]) not in seen and not seen.add((c[0], c[1]))]
    for planes, th, pf in cands:
        try:
            dev.SetTuning(planes, th, pf)
        except RuntimeError:
            continue
        dev.StepProfiled(dt, warmup)
        ms, _ = dev.StepProfiled(dt, steps)
        dev.Step(dt, 2 * steps)
        dev.Synchronize()
        walls = []
        for _ in range(3):
            t0 = time.perf_counter()
            dev.Step(dt, 2 * steps)
            dev.Synchronize()
            walls.append(time.perf_counter() - t0)
        step_ms = min(walls) * 1e3 / (2 * steps)
        res.append({"planes": planes, "threads": th, "prefetch": pf, "sweep_ms": ms / steps, "step_ms": step_ms})
        if log:
            log("tune planes=%d threads=%d prefetch=%d: sweep %.4f ms, step %.4f ms per step"
                % (planes, th, pf, ms / steps, step_ms))
    best = min(res, key=lambda r: r["step_ms"])
    dev.SetTuning(best["planes"], best["threads"], best["prefetch"])
    if save:
        c = load_cache()
        c[key(spec, dev.StepsPerPass())] = dict(best, all=res)
        with open(CACHE, "w") as f:
            json.dump(c, f, indent=1, sort_keys=True)
    return best, res
