"""Loading of the golden fixtures recorded from the reference (tests/golden/)."""
import glob
import hashlib
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")

RELAX = {"gs": 1, "jacobi": 0}
RESTR = {"inj": 0, "fw": 1}


def names(prefix):
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, prefix + "*.npz")))


def load(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    rec = {k: z[k] for k in z.files}
    rec["params"] = json.loads(str(rec["params"]))
    return rec


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def solve_kwargs(p):
    """Solver-construction arguments (oracle_lib.Solver / mgp.Solver) of a solve fixture."""
    dims = (int(p["nx"]), int(p["ny"]), int(p["nz"]))
    lengthx = float(p["lengthx"])
    return dict(dims=dims, levels=int(p["levels"]), h=lengthx / (dims[0] - 1.0),
                origin=(float(p["ox"]), float(p["oy"]), float(p["oz"])), bc=p["bc"],
                relaxation=RELAX[p["relaxation"]], restriction=RESTR[p["restriction"]],
                problem=-1 if int(p.get("init_zero", 0)) else int(p["problem"]))


def run_kwargs(p):
    return dict(cycle=p["cycle"], nsmooth=int(p["nsmooth"]), omega=float(p["omega"]),
                ncycles=int(p["ncycles"]))
