"""Regenerate arkus.json, the Arkus cluster table, from owlas/magpy's
magpy/geometry/arkus.py.

The table is published data (N. Arkus, V. N. Manoharan, M. P. Brenner,
Phys. Rev. Lett. 103, 118303 (2009)) in the reference's ordering and
orientation; configuration ids must select the same coordinates, so it is
kept identical.  The reference file is parsed as text with `ast` and only
its numeric literals are read; nothing from it is executed.

  python tests/golden/make_arkus_table.py /root/reference/magpy/geometry/arkus.py

It writes magpy_amd/geometry/arkus.json; the shipped package only reads that
file.
"""
import ast
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
OUT = os.path.join(ROOT, "magpy_amd", "geometry", "arkus.json")


def extract(path):
    tree = ast.parse(open(path).read())
    for node in tree.body:
        if isinstance(node, ast.Assign) and getattr(node.targets[0], "id", "") == "ARKUS":
            table = {}
            for k, v in zip(node.value.keys, node.value.values):
                table[str(ast.literal_eval(k))] = {
                    str(ast.literal_eval(kk)): ast.literal_eval(vv.args[0])
                    for kk, vv in zip(v.keys, v.values)}
            return table
    raise SystemExit("no ARKUS table in " + path)


if __name__ == "__main__":
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/magpy/geometry/arkus.py"
    with open(OUT, "w") as f:
        json.dump(extract(src), f, indent=0)
    print("wrote", OUT)
