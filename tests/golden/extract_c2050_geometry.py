"""Write tests/golden/geometry_c2050.json from the reference's sweep results.

The reference commits its Tesla C2050 sweep output as CSV data
(scripts/c2050/<stencil>.dat, produced by scripts/utils.py:36-244 from the
drivers' "Key: value" lines) next to the sweep definitions
(scripts/c2050/<stencil>.yml). Every row records the driver inputs (block
sizes, elements per thread, time tile, time steps; the problem size is the
yml's problem_size) together with the geometry the driver derived
(GeneratorParams::computeDerived) and the per-block counts it printed
(ProgramGenerator::printProgramParameters or the 3D drivers' own block).
Those are known-answer vectors for ot_derive in reference geometry.

This script only reads those data files as text; run it in a container that
has the reference mounted:  python tests/golden/extract_c2050_geometry.py
"""
import csv
import json
import os
import re

REF = "/root/reference/scripts/c2050"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "geometry_c2050.json")

FILES = {  # dat file -> engine stencil name
    "j2d": "jacobi-2d", "j3d": "jacobi-3d", "p2d": "poisson-2d", "p3d": "poisson-3d",
    "g2d": "gradient-2d", "g3d": "gradient-3d", "fdtd2d": "fdtd-2d",
    "rician2d": "rician-denoise-2d", "tv2d": "cdsc-tv-update-2d",
}
INPUTS = ["Block Size X", "Block Size Y", "Block Size Z", "Elements/Thread", "Time Tile Size",
          "Time Steps"]
OUTPUTS = ["Num Blocks X", "Num Blocks Y", "Num Blocks Z", "Padded Size", "Real Size",
           "Shared Size", "Real Per Block X", "Real Per Block Y", "Total FP/Block", "Total FP",
           "Useful FP/Block", "Useful FP", "Useful Ratio", "Global Loads/Block",
           "Global Stores/Block", "Shared Loads/Block", "Shared Stores/Block"]
MAX_ROWS = 80


def problem_size(yml_path):
    with open(yml_path) as fh:
        m = re.search(r"problem_size:\s*(\d+)", fh.read())
    return int(m.group(1))


def main():
    out = []
    for tag, name in FILES.items():
        n = problem_size(os.path.join(REF, tag + ".yml"))
        with open(os.path.join(REF, tag + ".dat")) as fh:
            rows = [r for r in csv.DictReader(fh)]
        rows = [r for r in rows if r.get("phase_limit", "0") in ("0", "")]
        # deterministic spread over the sweep, always keeping every time-tile size
        picked, seen_t = [], set()
        stride = max(1, len(rows) // MAX_ROWS)
        for i, r in enumerate(rows):
            t = r["Time Tile Size"]
            if i % stride == 0 or t not in seen_t:
                picked.append(r)
                seen_t.add(t)
        for r in picked:
            rec = {"stencil": name, "problem_size": n}
            for k in INPUTS:
                rec[k] = int(r[k])
            for k in OUTPUTS:
                if k in r and r[k] != "":
                    v = float(r[k])
                    rec[k] = int(v) if k != "Useful Ratio" and v == int(v) else v
            out.append(rec)
    with open(OUT, "w") as fh:
        json.dump(out, fh, indent=0)
    print(f"wrote {len(out)} vectors to {OUT}")


if __name__ == "__main__":
    main()
