"""Writes tests/golden/nd2fe14b.toml: the data of the reference's Nd2Fe14B setup
file (bench/assets/nd2fe14b.toml — 68-atom cell, nine sublattice species,
[interaction] and [symmetry] defaults) in the same TOML layout, without its
prose comments.  Run once in a tree that holds the reference:

    python tests/golden/make_setup_fixture.py /root/reference/bench/assets/nd2fe14b.toml
"""
import math
import os
import sys

import tomli


def _num(x):
    if isinstance(x, float) and math.isinf(x):
        return "inf" if x > 0 else "-inf"
    return repr(x)


def main(src):
    with open(src, "rb") as f:
        d = tomli.load(f)
    s, i, y = d["structure"], d["interaction"], d["symmetry"]
    out = ["[structure]", "lattice = ["]
    out += [f"  [{', '.join(_num(float(v)) for v in row)}]," for row in s["lattice"]]
    out += ["]", "positions = ["]
    out += [f"  [{', '.join(f'{float(v):.15f}' for v in row)}]," for row in s["positions"]]
    out += ["]", f"species = [{', '.join(str(int(v)) for v in s['species'])}]",
            f"species_labels = [{', '.join(repr(str(v)).replace(chr(39), chr(34)) for v in s['species_labels'])}]",
            "", "[interaction]", f"nbody = {int(i['nbody'])}", f"cutoff = {_num(float(i['cutoff']))}",
            f"lmax = [{', '.join(str(int(v)) for v in i['lmax'])}]",
            f"isotropy = {'true' if i['isotropy'] else 'false'}", "", "[symmetry]",
            f"backend = \"{y['backend']}\"", f"tol = {_num(float(y['tol']))}", ""]
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "nd2fe14b.toml")
    with open(dst, "w") as f:
        f.write("\n".join(out))
    with open(dst, "rb") as f:
        assert tomli.load(f) == d
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1])
