"""Loads the committed golden fixtures (tests/golden/*.npz) into workload objects."""
import os

import numpy as np

from svfsiplus_amd import workload as W

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["pipe_n2", "pipe_n4"]


def load(name):
    d = dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))
    p = W.Pipe(int(d["n"]), int(d["nz"]), d["x"], d["IEN"], np.zeros(d["x"].shape[0], np.int32), d["rowPtr"],
               d["colPtr"], 0.3, 3.0)
    p.faces = [("inlet", d["inlet"]), ("wall", d["wall"])]
    rho, mu, dt, ro_inf, peak = d["phys"].tolist()
    ph = W.Physics(rho=rho, mu=mu, dt=dt, ro_inf=ro_inf, peak=peak)
    return p, ph, d
