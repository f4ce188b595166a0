"""Extract data fixtures from the reference repository (run in the build
container only; the GPU hosts receive the committed outputs).

Writes:
  cc-pvdz-HCNO.dat      -- the H, C, N, O blocks of src/hf/data/basis/cc-pvdz.dat
  molecules/*.xyz       -- the reference geometries of the H/C/N/O molecules
  molecule_test.json    -- src/hf/benchmark/molecule_test.json (energies, wall
                           times) as published by the reference
"""
import json
import os
import re
import shutil

REF = "/root/reference/src/hf"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    text = open(os.path.join(REF, "data/basis/cc-pvdz.dat")).read()
    keep = []
    for block in text.split("****"):
        b = block.strip()
        if re.match(r"^(H|C|N|O)\s+0", b):
            keep.append(b)
    with open(os.path.join(HERE, "cc-pvdz-HCNO.dat"), "w") as fh:
        fh.write("\n****\n".join(keep) + "\n****\n")
    os.makedirs(os.path.join(HERE, "molecules"), exist_ok=True)
    for name in sorted(os.listdir(os.path.join(REF, "data/molecules"))):
        src = os.path.join(REF, "data/molecules", name)
        lines = open(src).read().strip().split("\n")
        syms = {ln.split()[0] for ln in lines[2:] if ln.strip()}
        if syms <= {"H", "C", "N", "O"}:
            shutil.copy(src, os.path.join(HERE, "molecules", name))
    data = json.load(open(os.path.join(REF, "benchmark/molecule_test.json")))
    with open(os.path.join(HERE, "molecule_test.json"), "w") as fh:
        json.dump(data, fh, indent=1)


if __name__ == "__main__":
    main()
