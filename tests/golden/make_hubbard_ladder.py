"""Known answers of the reference's Hubbard-ladder regression test as a JSON fixture.

Source (data files of the reference's own tests, read with the HDF5 command-line dump tool):
  dmrg/regression/applications/dmrg/ref/test_fermiladder.diag.h5  exact diagonalisation: energy,
      one-body density matrices of both spins (labels "( x,y ) -- ( x',y' )")
  dmrg/regression/applications/dmrg/ref/test_coded_fermiladder_ts.h5  two-site DMRG (TwoU1,
      M = 400): energy per sweep
The model (test_coded_fermiladder_ts.py): ALPS "open ladder" L = 6 (12 sites), fermion Hubbard,
t = t' = 1, U = 8, N_up = N_down = 10.  Writes hubbard_ladder.json next to this script.
"""
import json
import os
import re
import subprocess

REF = "/root/reference/dmrg/regression/applications/dmrg/ref"
H5DUMP = os.environ.get("H5DUMP", "/opt/conda/bin/h5dump")


def dump(path, dataset):
    out = subprocess.run([H5DUMP, "-m", "%.17g", "-d", dataset, path], capture_output=True, text=True,
                         check=True).stdout
    body = out[out.index("DATA {") + 6:]
    if "H5T_STRING" in out:
        return re.findall(r'"([^"]*)"', body)
    vals = re.findall(r"\):\s*([^\n]*)", body)
    return [float(x) for v in vals for x in v.replace(",", " ").split()]


def main():
    diag = os.path.join(REF, "test_fermiladder.diag.h5")
    base = "/spectrum/sectors/0/results/"
    out = {"model": {"lattice": "open ladder", "L": 6, "sites": 12, "t": 1.0, "t_rung": 1.0, "U": 8.0,
                     "n_up": 10, "n_dn": 10,
                     "site_of": "(x, y) -> 2 x + y (rung partners adjacent, ALPS ladder vertex order)"},
           "ed_energy": dump(diag, base + "Energy/mean/value")[0]}
    for spin in ("up", "down"):
        labels = dump(diag, base + "Onebody density matrix %s/labels" % spin)
        vals = dump(diag, base + "Onebody density matrix %s/mean/value" % spin)
        pairs = []
        for lab, v in zip(labels, vals):
            a, b = re.findall(r"\(\s*(\d+)\s*,\s*(\d+)\s*\)", lab)
            pairs.append([2 * int(a[0]) + int(a[1]), 2 * int(b[0]) + int(b[1]), v])
        out["ed_onebody_" + spin] = pairs
        out["ed_local_density_" + spin] = dump(diag, base + "Local density %s/mean/value" % spin)
    ts = os.path.join(REF, "test_coded_fermiladder_ts.h5")
    out["dmrg_ts_energy_per_sweep"] = [dump(ts, "/simulation/sweep%d/results/Energy/mean/value" % k)
                                       for k in range(2)]
    json.dump(out, open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "hubbard_ladder.json"), "w"),
              indent=1)
    print("ED energy", out["ed_energy"])


if __name__ == "__main__":
    main()
