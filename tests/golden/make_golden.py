"""Write the golden fixtures under tests/golden/ from the reference's own test data.

Run in a tree where the reference checkout is readable (default
/root/reference, override with QUIP_REF); the outputs are committed so the
tests never need the reference at run time.  Only data is copied: the inputs
and expected outputs that the reference's tests hold.

  si_soap_cur_points.npz   tests/Si.soap.cur_points.json (100 SOAP sparse points,
                           n_max=8 l_max=8 cutoff=4 ctw=1 atom_sigma=0.5) with
                           the 1-based atom indices they belong to, and every
                           frame of tests/Si.np1.xyz (positions, lattices)
  c2h_soap_grad.npz        tests/test_descriptor.py:61-190 (C2H structure,
                           grad_index_0based, grad_data[:2] reference values)
  gap_water/               tests/GAP.xml + 3 sparseX files + tests/gap_sample.xyz
                           (distance_2b ard_se model, E and F reference)
  potential_cell.json      tests/test_potential_cell.py:30-44 (5 small-cell energies)
"""
import ast
import json
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(os.environ.get("QUIP_REF", "/root/reference"), "tests")
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from quip_amd.atoms import read_xyz  # noqa: E402


def main():
    d = json.load(open(os.path.join(REF, "Si.soap.cur_points.json")))
    X = np.array(d["coords"][0]["sparsex"]).reshape(100, 325)
    frames = read_xyz(os.path.join(REF, "Si.np1.xyz"))
    nat = np.array([len(f) for f in frames])
    np.savez_compressed(os.path.join(HERE, "si_soap_cur_points.npz"), sparsex=X,
                        index=np.array(d["index"][0]), config=np.array(d["config"]),
                        natoms=nat, positions=np.vstack([f.positions for f in frames]),
                        lattices=np.array([f.cell for f in frames]),
                        numbers=np.concatenate([f.numbers for f in frames]))

    src = open(os.path.join(REF, "test_descriptor.py")).read()
    grad = np.array(ast.literal_eval(src.split("self.ref_grad_array = np.array(")[1].split(")\n")[0]))
    gidx = np.array(ast.literal_eval(src.split("self.ref_grad_index_0based = np.array(")[1].split(", dtype")[0]))
    np.savez_compressed(os.path.join(HERE, "c2h_soap_grad.npz"), grad=grad, grad_index=gidx,
                        positions=np.array([[0., 0., 0.], [0.875, 0.875, 0.875], [0.2, 0.2, 0.1]]),
                        numbers=np.array([6, 6, 1]),
                        cell=np.array([[0.0, 1.75, 1.75], [1.75, 0.0, 1.75], [1.75, 1.75, 0.0]]),
                        descriptor=np.array("soap cutoff=1.3 l_max=4 n_max=4 atom_sigma=0.5 n_Z=2 Z={1 6}"),
                        nl_cutoff=3.0)

    out = os.path.join(HERE, "gap_water")
    os.makedirs(out, exist_ok=True)
    for fn in os.listdir(REF):
        if fn.startswith("GAP.xml") or fn == "gap_sample.xyz":
            shutil.copy(os.path.join(REF, fn), os.path.join(out, fn))

    src = open(os.path.join(REF, "test_potential_cell.py")).read()
    energies = ast.literal_eval("[" + src.split("self.ref_energies = [")[1].split("]")[0] + "]")
    json.dump({"cell_sizes": list(np.linspace(2.5, 4.5, 5)), "positions": [[0, 0, 0], [1, 1, 1]],
               "numbers": [1, 1], "energies": energies, "tolerance": 1e-6},
              open(os.path.join(HERE, "potential_cell.json"), "w"), indent=1)


if __name__ == "__main__":
    main()
