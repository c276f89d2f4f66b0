"""Writes the golden fixtures of this directory from the FP64 oracle and the
exact (closed-form) solvers. The reference ships no stored vectors (its tests
use unseeded randn), so these are produced by the oracle, whose own pins are
in tests/test_oracle.py. Run: python tests/golden/make_golden.py"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import catn_oracle as O  # noqa: E402
from oracle import configs as Cf  # noqa: E402
from oracle import exact as X  # noqa: E402


def run(name, n, edges, w, h, beta, Dmax, chi, exact=None):
    tn = O.ising_network(n, edges, w, h, beta, Dmax=Dmax, chi=chi)
    s = O.Schedule()
    lnZ, err, psi = O.contraction(tn, s)
    cuts = [e for e in s.events if e[0] == "cut_keep"]
    rec = {
        "name": name, "n": n, "beta": beta, "Dmax": Dmax, "chi": chi,
        "edges_in": [list(e) for e in edges], "w": list(map(float, w)), "h": list(map(float, h)),
        "lnZ": lnZ, "lnZ_per_N": lnZ / n, "error": err, "psi": psi,
        "edge_sequence": [list(e) for e in s.edges],
        "cut_kept": [c[2] for c in cuts],
        "maxdim_intermediate": tn.maxdim_intermediate,
        "exact_lnZ": exact,
    }
    with open(os.path.join(HERE, name + ".json"), "w") as f:
        json.dump(rec, f)
    print(name, lnZ / n, exact / n if exact is not None else None)


def main():
    n, e, w, h = Cf.ising_lattice(16)
    run("config1_ising16", n, e, w, h, 0.44, 20, 200, X.lattice_lnZ(16, w, h, 0.44))
    n, e, w, h = Cf.spin_glass_3reg(100, seed=0)
    run("config2_sg100", n, e, w, h, 1.0, 32, 256, X.elimination_lnZ(n, e, w, h, 1.0))
    n, e, w, h = Cf.spin_glass_3reg(20, seed=3)
    run("sg20", n, e, w, h, 0.8, 8, 16, X.brute_lnZ(n, e, w, h, 0.8))
    n, e, w, h = Cf.ising_lattice(4)
    beta = float(np.log(1 + np.sqrt(2)) / 2)
    run("ising4_exact", n, e, w, h, beta, -1, 100000, X.brute_lnZ(n, e, w, h, beta))


if __name__ == "__main__":
    main()
