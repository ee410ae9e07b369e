"""Writes the reference-held known-answer fixtures used by the oracle tests.

Run in a container where the acoustic_petsc tree is mounted at
/root/reference.  It copies data only (numbers from result files), never
source:
  * 3d/weno4/plane_wave/results/main.tex:37-42  -> plane_wave_3d_table.json
  * 2d/weno4/standing_wave/results/Error.dat    -> standing_wave_2d_error.json
  * 3d/ctcs/point_source/results/p.dat          -> ctcs_point_source_p.json
"""
import json
import os
import re

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rows = []
    with open(os.path.join(REF, "3d/weno4/plane_wave/results/main.tex")) as f:
        for line in f:
            m = re.match(r"\s*(\d+)\s*&\s*([0-9.e+-]+)\s*&\s*([0-9.e+-]+)\s*&", line)
            if m:
                rows.append({"N": int(m.group(1)), "l2": float(m.group(2)), "linf": float(m.group(3)),
                             "l2_text": m.group(2), "linf_text": m.group(3)})
    with open(os.path.join(HERE, "plane_wave_3d_table.json"), "w") as f:
        json.dump({"source": "3d/weno4/plane_wave/results/main.tex:37-42", "rows": rows}, f, indent=1)

    rows = []
    with open(os.path.join(REF, "2d/weno4/standing_wave/results/Error.dat")) as f:
        for line in f:
            p = line.split()
            if len(p) == 3:
                rows.append({"N": int(p[0]), "l2": float(p[1]), "linf": float(p[2]),
                             "l2_text": p[1], "linf_text": p[2]})
    with open(os.path.join(HERE, "standing_wave_2d_error.json"), "w") as f:
        json.dump({"source": "2d/weno4/standing_wave/results/Error.dat:1-4", "rows": rows}, f, indent=1)

    t, p = [], []
    with open(os.path.join(REF, "3d/ctcs/point_source/results/p.dat")) as f:
        for line in f:
            a = line.split()
            if len(a) == 2:
                t.append(a[0])
                p.append(a[1])
    with open(os.path.join(HERE, "ctcs_point_source_p.json"), "w") as f:
        json.dump({"source": "3d/ctcs/point_source/results/p.dat:1-251", "t_text": t, "p_text": p}, f)


if __name__ == "__main__":
    main()
