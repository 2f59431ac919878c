"""Write tests/golden/force_cases.npz: small force-field cases evaluated by the CPU oracle.

Each case stores a system (positions, domain vectors, grid, topology tables, per-type
parameters) and the oracle's energy, gradient and virial for both nanocell models, computed
with the literal per-cell loop of oracle.mmff (a restatement of micmec/pes/mmff.py:326-403).
The oracle itself is pinned by tests/test_oracle_force.py; these vectors freeze its output so
that later changes of either side are caught. Run from the repository root:

    python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from micmec_amd import units  # noqa: E402
from micmec_amd.system import System  # noqa: E402
from micmec_amd.workloads import make_system  # noqa: E402
from oracle import mmff as omm  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "force_cases.npz")
CHK = os.path.join(ROOT, "tests", "golden", "chk")


def system_arrays(prefix, s):
    d = {}
    d[prefix + "pos"] = np.asarray(s.pos, dtype=np.float64)
    d[prefix + "rvecs"] = np.asarray(s.domain.rvecs, dtype=np.float64)
    d[prefix + "masses"] = np.asarray(s.masses, dtype=np.float64)
    d[prefix + "grid"] = np.rint(np.asarray(s.grid, dtype=np.float64)).astype(np.int64)
    d[prefix + "types"] = np.rint(np.asarray(s.types, dtype=np.float64)).astype(np.int64)
    d[prefix + "surrounding_nodes"] = np.asarray(s.surrounding_nodes, dtype=np.int64)
    d[prefix + "surrounding_cells"] = np.asarray(s.surrounding_cells, dtype=np.int64)
    d[prefix + "boundary_nodes"] = np.asarray(s.boundary_nodes, dtype=np.int64)
    for key, value in s.params.items():
        if key.split("/")[-1] in ("cell", "elasticity", "free_energy", "effective_temp", "mass"):
            d[prefix + "param:" + key] = np.asarray(value, dtype=np.float64)
    return d


def load_system(npz, prefix):
    params = {}
    for key in npz.files:
        if key.startswith(prefix + "param:"):
            v = npz[key]
            params[key[len(prefix) + 6:]] = float(v) if v.ndim == 0 else v
    return System(pos=npz[prefix + "pos"].copy(), masses=npz[prefix + "masses"].copy(),
                  rvecs=npz[prefix + "rvecs"].copy(),
                  surrounding_cells=npz[prefix + "surrounding_cells"],
                  surrounding_nodes=npz[prefix + "surrounding_nodes"],
                  boundary_nodes=npz[prefix + "boundary_nodes"], grid=npz[prefix + "grid"],
                  types=npz[prefix + "types"], params=params)


def cases():
    out = []
    out.append(("nve3", make_system((3, 3, 3), seed=0)))
    out.append(("mix4_hollow", make_system((4, 4, 4), p_b=0.5, p_hole=0.2, seed=5,
                                           noise=0.1 * units.angstrom)))
    s = System.from_file(os.path.join(CHK, "3x3x3_fcu_hollow_micmec.chk"))
    s.pos = s.pos + np.random.default_rng(3).normal(0, 0.1, s.pos.shape)
    out.append(("fcu_hollow3", s))
    s = System.from_file(os.path.join(CHK, "2x2x2_reo_micmec.chk"))
    s.pos = s.pos + np.random.default_rng(4).normal(0, 0.1, s.pos.shape)
    out.append(("reo2_nondiag_h0", s))
    out.append(("finite_3x3x2", make_system((3, 3, 2), p_b=0.4, seed=6, pbc=False,
                                            noise=0.1 * units.angstrom)))
    out.append(("mix2_strained", make_system((2, 3, 2), p_b=0.5, seed=8, strain=0.03,
                                             noise=0.1 * units.angstrom)))
    return out


def main():
    data = {}
    names = []
    for name, s in cases():
        names.append(name)
        prefix = name + "/"
        data.update(system_arrays(prefix, s))
        sysd = {"pos": s.pos, "rvecs": s.domain.rvecs, "grid": s.grid, "types": s.types,
                "surrounding_nodes": s.surrounding_nodes,
                "surrounding_cells": s.surrounding_cells, "params": s.params}
        for model in ("corner8", "original"):
            e, g, v = omm.compute_loop(sysd, model)
            data[prefix + model + ":energy"] = np.float64(e)
            data[prefix + model + ":gpos"] = g
            data[prefix + model + ":vtens"] = v
            print("%-18s %-9s E = %.15e" % (name, model, e))
    data["cases"] = np.array(names)
    np.savez_compressed(OUT, **data)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
