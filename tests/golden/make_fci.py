"""Golden FCI energies of the seeded quantum-chemistry problems (determinant CI in
qcmaquis_amd.qc.fci_energy, itself pinned against the dense MPO spectrum for small L).
Writes fci_golden.json next to this script.

config 1: L = 8 orbitals, 8 electrons, S_z = 0, integrals make_integrals(8, seed=1001, rank=16);
small:    L = 6, 6 electrons, seed 7, rank 8 (fast GPU sweep test).
"""
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))

from qcmaquis_amd.qc import fci_energy, make_integrals  # noqa: E402


def main():
    out = {}
    for name, L, seed, rank, nu, nd in [("config1", 8, 1001, 16, 4, 4), ("small6", 6, 7, 8, 3, 3)]:
        h, eri, ec = make_integrals(L, seed, rank)
        e = fci_energy(h, eri, nu, nd, ecore=ec, k=2)
        out[name] = dict(L=L, seed=seed, rank=rank, n_up=nu, n_dn=nd, e0=float(e[0]), e1=float(e[1]))
        print(name, e)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fci_golden.json")
    json.dump(out, open(path, "w"), indent=1)


if __name__ == "__main__":
    main()
