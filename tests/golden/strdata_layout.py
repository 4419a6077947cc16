"""Writes strdata_layout.json: the byte offset and size of every field of the reference's
host structs (strdata.h:60-540: ND, FC, LEAF, TREE, BRANCH, CELL, FOREST, PAR, CON, SOL,
RUN) and each struct's size, on x86-64 with gcc and the MPICH mpi.h of this image
(MPI_Comm is an int there).  Only the struct declarations are compiled: they need
nothing from the METIS and PETSc headers that strdata.h also includes.

    python tests/golden/strdata_layout.py [/root/reference/strdata.h]

include/fv_strdata.h must reproduce this layout (tests/test_abi_layout.py), so that a RUN
built by the reference's host code can be handed to the library.
"""
from __future__ import annotations

import json
import os
import pathlib
import re
import subprocess
import sys
import tempfile

STRUCTS = ["ND", "FC", "LEAF", "TREE", "BRANCH", "CELL", "FOREST", "PAR", "CON", "SOL", "RUN"]
HERE = pathlib.Path(__file__).resolve().parent
MPI_INCLUDE = "/opt/conda/include"


def struct_bodies(text: str) -> dict[str, str]:
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    out = {}
    for m in re.finditer(r"typedef\s+struct\s+(\w+)\s*\{(.*?)\}\s*(\w+)\s*;", text, flags=re.S):
        if m.group(3) in STRUCTS:
            out[m.group(3)] = m.group(2)
    return out


def field_names(body: str) -> list[str]:
    names = []
    for decl in body.split(";"):
        decl = " ".join(decl.split())
        if not decl:
            continue
        # the base type ends at the first declarator; declarators are comma-separated
        first, *rest = decl.split(",")
        names.append(re.findall(r"(\w+)\s*(?:\[[^\]]*\]\s*)*$", first)[0])
        for d in rest:
            names.append(re.findall(r"(\w+)\s*(?:\[[^\]]*\]\s*)*$", d)[0])
    return names


def probe_program(prelude: str, fields: dict[str, list[str]]) -> str:
    lines = [prelude, "#include <stddef.h>", "#include <stdio.h>", "int main(void) {"]
    for s, names in fields.items():
        lines.append(f'  printf("{s} %zu\\n", sizeof({s}));')
        for n in names:
            lines.append(f'  printf("{s}.{n} %zu %zu\\n", offsetof({s}, {n}), sizeof((({s} *)0)->{n}));')
    lines += ["  return 0;", "}"]
    return "\n".join(lines) + "\n"


def run_probe(src: str, includes: list[str]) -> dict:
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-std=gnu11", *[f"-I{i}" for i in includes], c, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    res = {}
    for line in out.splitlines():
        k, *v = line.split()
        res[k] = [int(x) for x in v]
    return res


def reference_layout(header: str) -> tuple[dict, dict]:
    bodies = struct_bodies(open(header).read())
    fields = {s: field_names(bodies[s]) for s in STRUCTS}
    decls = "\n".join(f"typedef struct {s} {s};" for s in STRUCTS)
    defs = "\n".join(f"struct {s} {{{bodies[s]}}};" for s in STRUCTS)
    src = probe_program("#include <mpi.h>\n" + decls + "\n" + defs, fields)
    return fields, run_probe(src, [MPI_INCLUDE])


def main():
    header = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/strdata.h"
    fields, layout = reference_layout(header)
    json.dump({"source": "strdata.h:60-540 (reference), gcc x86-64, MPICH mpi.h (MPI_Comm = int)",
               "fields": fields, "layout": layout}, open(HERE / "strdata_layout.json", "w"), indent=0)
    print(f"{sum(map(len, fields.values()))} fields of {len(fields)} structs")


if __name__ == "__main__":
    main()
