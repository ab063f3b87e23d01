"""The reference's stationary mouth model for quasi-stationary release sites.

QuasyStationaryReleaseSite (QuasyStationaryReleaseSite.cpp:10-50) takes the
mouth [Ca2+] of an open set from a StationaryDiffusionModel: background plus,
over the open channels, the [Ca2+] each adds at the others' mouths
(StationaryDiffusionModel.cpp:3-11).  Its one implementation,
LinearNaraghiNeher (LinearNaraghiNeher.cpp:9-75), fills that table from the
linearised steady state of a point source with mobile buffers: a sum over
the eigenmodes of B = -Dm^-1 AA of coef_m exp(-r decay_m) / r, and the
"Open Concentration" at a channel's own mouth.

The modes are computed here on the host (numpy's eigensolver in place of
Eigen's); the device path (casim_qs_set_stationary_model) evaluates the sum
per channel pair in FP64 at every event.  The mouth values do not depend on
the jSR, as in the reference.

Conditioning: the conserved mode (total Ca) has eigenvalue zero only to
rounding, about eps * max|mu|, and the model takes its square root, so two
eigensolvers (Eigen's, numpy's) can differ by sqrt(eps * max|mu|) * 1e-8 in
that mode's decay; added(r) then moves by up to that times the mode's
coefficient at every r (about 1e-8 uM for buffers with max|mu| ~ 1e12).
"""
import dataclasses
import json
import math
import os

import numpy as np

FARADAY = 96485.33e-6  # StationaryDiffusionModel.h: Far


@dataclasses.dataclass
class StationaryModel:
    """added(r) = sum_m coef[m] exp(-r decay[m]) / r for r > 1e-6, ca_open at
    r <= 1e-6 (a channel's own mouth); the mouth [Ca2+] of channel j is
    background + sum over open i of added(|p_i - p_j|)."""
    coef: np.ndarray
    decay: np.ndarray
    background: float
    ca_open: float

    def added(self, r):
        if not r > 1e-6:
            return self.ca_open
        v = 0.0
        for c, d in zip(self.coef, self.decay):
            v += 1.0 / r * c * math.exp(-r * d)
        return v


def linear_naraghi_neher(buffers, current, ca0, dca, ca_open):
    """LinearNaraghiNeher's modes.  buffers: {name: {"kon", "koff", "D",
    "Total Concentration"}} (kon, koff, D and the concentrations in the
    units of the release-site file); current: the channel's "Current";
    ca0, dca: the cleft's Ca concentration and D; ca_open: "Open
    Concentration".  Every buffer must be mobile: the reference inverts
    diag(D)."""
    names = sorted(buffers)  # the reference iterates a std::map
    n = len(names)
    dm = np.empty(n + 1)
    rates = np.zeros((n + 1, n + 1))  # d/dt of (bound_1 .. bound_n, Ca), linearised at ca0
    for i, name in enumerate(names):
        b = buffers[name]
        if not b["D"] > 0:
            raise ValueError("LinearNaraghiNeher needs every buffer mobile (D > 0): %r has D = %r" % (name, b["D"]))
        K = b["koff"] / b["kon"]
        slope = b["Total Concentration"] * K / ((K + ca0) * (K + ca0))  # d(bound)/d(Ca) at equilibrium
        relax = b["koff"] + b["kon"] * ca0                            # 1 / tau
        rates[i, i] = -relax
        rates[i, n] = slope * relax
        rates[n, i] = relax
        rates[n, n] -= slope * relax
        dm[i] = b["D"]
    if not dca > 0:
        raise ValueError("the Ca diffusion coefficient must be > 0")
    dm[n] = dca
    B = -rates / dm[:, None]
    w, v = np.linalg.eig(B)
    mu, V = np.real(w), np.real(v)
    source = np.zeros(n + 1)
    source[n] = current / (4 * math.pi * FARADAY * dca)
    amp = np.linalg.solve(V, source)
    # the Ca row of every mode, with the reference's 1e-4 (concentration) and
    # 1e-8 (length) unit factors folded in
    coef = amp * V[n, :] * 1e-4
    decay = np.sqrt(np.abs(mu)) * 1e-8
    return StationaryModel(coef=np.ascontiguousarray(coef), decay=np.ascontiguousarray(decay),
                           background=float(ca0), ca_open=float(ca_open))


def from_release_site_json(path_or_dict, base_dir=None):
    """The "Calcium Release Site" block of a reference parameter file
    (QuasyStationaryReleaseSite.cpp:10-50): (model, channel name, kinetic
    model name, kinetic parameters, channel coordinates [n, 2] in the file's
    units, Max Time, the cleft's ions).  Only "LinearNaraghiNeher" exists."""
    if isinstance(path_or_dict, (str, os.PathLike)):
        with open(path_or_dict) as f:
            d = json.load(f)
        base_dir = base_dir or os.path.dirname(os.path.abspath(path_or_dict))
    else:
        d = path_or_dict
    crs = d["Calcium Release Site"]
    if crs["Name"] != "LinearNaraghiNeher":
        raise ValueError("unknown stationary diffusion model %r" % crs["Name"])
    par = crs["Parameters"]
    chans = par["Channels"]
    if len(chans) != 1:
        raise ValueError("stationary diffusion models accept exactly one ion channel")
    name = next(iter(chans))
    ch = chans[name]
    coords = os.path.join(base_dir or ".", ch["Coordinates"])
    xy = np.loadtxt(coords, dtype=np.float64, ndmin=2)[:, :2]
    cleft = par["Dyadic Cleft"]
    ca = cleft["Ions"]["Calcium"]
    model = linear_naraghi_neher(cleft.get("Buffers", {}), ch["Current"], ca["Concentration"], ca["D"],
                                 par["Open Concentration"])
    ions = {k: v["Concentration"] for k, v in cleft["Ions"].items()}
    return {"model": model, "channel": name, "kinetics": ch["Model"], "params": ch["Parameters"], "xy": xy,
            "t_max": crs["Max Time"], "ions": ions}
