"""Extract data fixtures from the reference repository (run in the build
container only; the GPU hosts receive the committed outputs).

Writes:
  cc-pvdz-HCNO.dat      -- the H, C, N, O blocks of src/hf/data/basis/cc-pvdz.dat
  cc-pvdz-HS.dat        -- its H and S blocks (sulfur: the element whose header
                           line the reference's parser misreads, molecule.py:128-133)
  cc-pvtz-HCO.dat       -- the H, C, O blocks of src/hf/data/basis/cc-pvtz.dat
                           (f shells on C and O)
  molecules/*.xyz       -- the reference geometries of the H/C/N/O molecules and
                           of hydrogen_sulfide
  molecule_test.json    -- src/hf/benchmark/molecule_test.json (energies, wall
                           times) as published by the reference
  psi4_energies.json    -- "@RHF Final Energy" of every Psi4 output the reference
                           ships (src/hf/benchmark/psi4_output/*.out)
"""
import json
import os
import re
import shutil

REF = "/root/reference/src/hf"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    for src, out, elems in (("cc-pvdz.dat", "cc-pvdz-HCNO.dat", "H|C|N|O"),
                            ("cc-pvdz.dat", "cc-pvdz-HS.dat", "H|S"),
                            ("cc-pvtz.dat", "cc-pvtz-HCO.dat", "H|C|O")):
        text = open(os.path.join(REF, "data/basis", src)).read()
        keep = []
        for block in text.split("****"):
            b = block.strip()
            if re.match(rf"^({elems})\s+0", b):
                keep.append(b)
        with open(os.path.join(HERE, out), "w") as fh:
            fh.write("\n****\n".join(keep) + "\n****\n")
    os.makedirs(os.path.join(HERE, "molecules"), exist_ok=True)
    for name in sorted(os.listdir(os.path.join(REF, "data/molecules"))):
        src = os.path.join(REF, "data/molecules", name)
        lines = open(src).read().strip().split("\n")
        syms = {ln.split()[0] for ln in lines[2:] if ln.strip()}
        if syms <= {"H", "C", "N", "O"} or name == "hydrogen_sulfide.xyz":
            shutil.copy(src, os.path.join(HERE, "molecules", name))
    data = json.load(open(os.path.join(REF, "benchmark/molecule_test.json")))
    with open(os.path.join(HERE, "molecule_test.json"), "w") as fh:
        json.dump(data, fh, indent=1)
    psi4 = {}
    pdir = os.path.join(REF, "benchmark/psi4_output")
    for name in sorted(os.listdir(pdir)):
        m = re.search(r"@RHF Final Energy:\s+(-?[0-9.]+)", open(os.path.join(pdir, name)).read())
        if m:
            psi4[name[:-4]] = float(m.group(1))
    with open(os.path.join(HERE, "psi4_energies.json"), "w") as fh:
        json.dump(psi4, fh, indent=1)


if __name__ == "__main__":
    main()
