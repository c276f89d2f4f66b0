"""Writes config3_oracle.json: the FP64 oracle's contraction of all 64
instances of the batched-lattice configuration (oracle/configs.py
batch_instances(64, seed=2025): open L x L +-J lattices, L uniform in
{8, ..., 24}, h = 0, beta uniform in [0.2, 1.2]) at Dmax = 32, chi = 128:
lnZ, discarded weight, largest intermediate bond and the Dmax-cut kept
counts per instance. About 3 minutes on 8 cores.
Run: python tests/golden/make_config3.py"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import catn_oracle as O  # noqa: E402
from oracle import configs as Cf  # noqa: E402

SEED, COUNT, DMAX, CHI = 2025, 64, 32, 128


def main():
    out = {"seed": SEED, "count": COUNT, "Dmax": DMAX, "chi": CHI, "instances": []}
    for q, (L, n, edges, w, h, beta) in enumerate(Cf.batch_instances(COUNT, SEED)):
        t0 = time.time()
        tn = O.ising_network(n, edges, w, h, beta, Dmax=DMAX, chi=CHI)
        s = O.Schedule()
        lnZ, err, psi = O.contraction(tn, s)
        out["instances"].append({
            "index": q, "L": L, "beta": beta, "lnZ": lnZ, "lnZ_per_N": lnZ / n, "error": err,
            "psi": psi, "maxdim_intermediate": tn.maxdim_intermediate,
            "cut_kept": [e[2] for e in s.events if e[0] == "cut_keep"],
        })
        print(q, L, round(beta, 3), lnZ / n, f"{time.time() - t0:.1f} s", flush=True)
    with open(os.path.join(HERE, "config3_oracle.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()
