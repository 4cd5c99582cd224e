"""Writes tests/golden/helmholtz_fmm3d.npz from the reference's fixture
fmm-ref/fixtures/helmholtz_fmm3d.json (FMM3D v2.1.0 h3ddirectcg: potential and gradient
of e^{ik|x-y|}/(4 pi |x-y|), 200 sources, 40 targets, 12 cases of k and charges; values
at the sources with the self term dropped, then at the targets).  Run in the build
container only; the npz holds the same numbers as float64 arrays."""
import json
import os
import sys

import numpy as np

SRC = "/root/reference/fmm-ref/fixtures/helmholtz_fmm3d.json"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "helmholtz_fmm3d.npz")


def c(v):
    return complex(float(v[0]), float(v[1]))


def main():
    d = json.load(open(SRC))
    sources = np.array([[float(x) for x in p] for p in d["sources"]])
    targets = np.array([[float(x) for x in p] for p in d["targets"]])
    charges = {name: np.array([c(q) if isinstance(q, list) else complex(float(q), 0.0) for q in vals])
               for name, vals in d["charges"].items()}
    ks, kinds, pots, grads = [], [], [], []
    for case in d["cases"]:
        ks.append(c(case["k"]))
        kinds.append(0 if case["charges"] == "real" else 1)
        pots.append([c(v) for v in case["potential"]])
        grads.append([[complex(float(g[0]), float(g[1])), complex(float(g[2]), float(g[3])),
                       complex(float(g[4]), float(g[5]))] for g in case["gradient"]])
    np.savez_compressed(OUT, sources=sources, targets=targets, charges_real=charges["real"],
                        charges_complex=charges["complex"], k=np.array(ks), charge_kind=np.array(kinds),
                        potential=np.array(pots), gradient=np.array(grads))
    print(OUT, os.path.getsize(OUT))


if __name__ == "__main__":
    sys.exit(main())
