"""Writes the golden fixtures under tests/golden/ from the reference's own data files.

Run where the reference tree exists:  python tests/golden/make_golden.py
  clique_500_cp4.sol  -- the 2000 values of script/sol.txt:43-2043 ("ext value" per line,
                         in the reference's output order) together with its header facts
                         (sweeps 2011, MG_pesi 1992: script/sol.txt:31-32).
The input clique_500_cp4.txt.gz and the eg_fig2*.txt inputs are copied unchanged from
script/ and script/multi/.
"""
import os
import re
import shutil

REF = "/root/reference/script"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    vals = []
    for line in open(os.path.join(REF, "sol.txt")):
        mo = re.match(r"^(\d+)\)\t(-?\d+)$", line.rstrip("\n"))
        if mo:
            vals.append((int(mo.group(1)), int(mo.group(2))))
    with open(os.path.join(HERE, "clique_500_cp4.sol"), "w") as f:
        f.write("# from script/sol.txt: --eg0 --cpu, 2011 sweeps, MG_pesi=1992, 2000 vertices, 1746500 edges\n")
        for e, v in vals:
            f.write(f"{e} {v}\n")
    shutil.copy(os.path.join(REF, "clique_500_cp4.txt.gz"), HERE)
    for name in os.listdir(os.path.join(REF, "multi")):
        if name.startswith("eg_fig2") and name.endswith(".txt"):
            shutil.copy(os.path.join(REF, "multi", name), HERE)


if __name__ == "__main__":
    main()
