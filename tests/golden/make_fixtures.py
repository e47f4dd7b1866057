"""Regenerate the committed golden fixtures from the reference's test data.

Reads the reference's own test inputs (HDF5, dumped with the h5dump tool, no
reference code is imported or run) and golden output files, and writes:
  * GF_Chi_input.json, 4ring_input.json, 4ring_nomagfield_input.json
    (test/input/*/input.h5 as JSON: hopping, interaction, chemical potential,
    magnetic field, sectors, GF/Chi orbital lists);
  * gom1.dat, xiats.dat, xiatd.dat copied verbatim (test/input/GF_Chi/);
  * the three input.h5 files copied verbatim as <case>_input.h5 (read by the
    drop-in's HDF5 reader, whose tests compare them with the JSON above).
Usage: python tests/golden/make_fixtures.py [/root/reference]
"""
import json
import os
import re
import shutil
import subprocess
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))
H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"


def h5_to_dict(path):
    txt = subprocess.check_output([H5DUMP, path]).decode()
    out = {}
    stack = []
    lines = txt.splitlines()
    i = 0
    while i < len(lines):
        ln = lines[i].strip()
        m = re.match(r'GROUP "(.+)" \{', ln)
        if m:
            if m.group(1) != "/":
                stack.append(m.group(1))
            i += 1
            continue
        m = re.match(r'DATASET "(.+)" \{', ln)
        if m:
            name = "/".join(stack + [m.group(1)])
            shape = None
            data = []
            i += 1
            while not lines[i].strip().startswith("DATA {"):
                sm = re.search(r"SIMPLE \{ \( ([\d, ]+) \)", lines[i])
                if sm:
                    shape = [int(x) for x in sm.group(1).split(",")]
                if "SCALAR" in lines[i]:
                    shape = []
                i += 1
            i += 1
            while lines[i].strip() != "}":
                body = re.sub(r"\([\d,]+\):", "", lines[i])
                data += [float(x) for x in body.replace(",", " ").split()]
                i += 1
            out[name] = dict(shape=shape, data=data)
            i += 1
            continue
        if ln == "}" and stack:
            stack.pop()
        i += 1
    return out


for src, dst in (("GF_Chi", "GF_Chi_input.json"), ("4ring", "4ring_input.json"),
                 ("4ring_nomagfield", "4ring_nomagfield_input.json")):
    d = h5_to_dict(os.path.join(REF, "test", "input", src, "input.h5"))
    with open(os.path.join(OUT, dst), "w") as f:
        json.dump(d, f, indent=1, sort_keys=True)
for name in ("gom1.dat", "xiats.dat", "xiatd.dat"):
    shutil.copyfile(os.path.join(REF, "test", "input", "GF_Chi", name), os.path.join(OUT, name))
for case in ("4ring", "4ring_nomagfield", "GF_Chi"):
    shutil.copyfile(os.path.join(REF, "test", "input", case, "input.h5"), os.path.join(OUT, case + "_input.h5"))
print("fixtures written to", OUT)
