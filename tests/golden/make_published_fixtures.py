"""Extract the source-point prints of the reference's published benchmark logs
(test_20240512/result_*.txt in the reference tree) into published_traces.json.

Run in a container that has the reference tree; the JSON it writes is the
committed fixture (data only: the printed values, their step numbers and the
run headers).
"""
import glob
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "published_traces.json")


def parse(path):
    steps, vals, header = [], [], {}
    for line in open(path):
        m = re.match(r"TimeStep=(\d+)\s*;.*=\s*(\S+)\s*$", line)
        if m:
            steps.append(int(m.group(1)))
            vals.append(m.group(2))
            continue
        m = re.match(r"Number of grids: nx=(\d+), ny=(\d+), nz=(\d+)", line)
        if m:
            header["grid"] = [int(x) for x in m.groups()]
        m = re.match(r"ex, ey, ez=(\d+), (\d+), (\d+)", line)
        if m:
            header["elements"] = [int(x) for x in m.groups()]
        m = re.match(r"nx, ny, nz=(\d+), (\d+), (\d+)", line)
        if m:
            header["nodes"] = [int(x) for x in m.groups()]
        m = re.match(r"Sample (\d+)\s*: sourceTerm = (\S+)", line)
        if m:
            header.setdefault("source_samples", []).append([int(m.group(1)), m.group(2)])
        m = re.match(r"Element number for the source location: (\d+)", line)
        if m:
            header["source_element"] = int(m.group(1))
    return {"steps": steps, "values": vals, "header": header}


def main():
    out = {}
    for p in sorted(glob.glob(os.path.join(REF, "test_20240512", "result_*.txt"))):
        out[os.path.basename(p)] = parse(p)
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", OUT, len(out), "logs")


if __name__ == "__main__":
    main()
