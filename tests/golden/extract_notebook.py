"""Writes notebook_runs.json: printed outputs of the reference's own runs.

Reads /root/reference/cgle.ipynb as JSON text (no code from it is run) and
keeps, as data: the factor-index listing of `rdcu -N 256,256 -n 2 -h`, the
`-v` equation lines for the 2-D CGLE coupling file, and for each DP45 run
(no -S) its command line, accepted steps and final h from the progress line.
Run from the repository root: python tests/golden/extract_notebook.py
"""
import json
import os
import re

NB = "/root/reference/cgle.ipynb"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "notebook_runs.json")


def main():
    nb = json.load(open(NB))
    res = {"source": "cgle.ipynb (znicolaou/rdcu)", "index_listing": None, "equations_2d": None,
           "runs": []}
    for i, c in enumerate(nb["cells"]):
        if c["cell_type"] != "code":
            continue
        for o in c.get("outputs", []):
            txt = "".join(o.get("text", []))
            if "Indices for 2 field(s) in 2 dimension(s):" in txt and res["index_listing"] is None:
                k = txt.index("Indices for")
                res["index_listing"] = txt[k:].rstrip("\n").split("\n")
                res["index_listing_cell"] = i
            for line in txt.split("\n"):
                if "./rdcu" in line and " -S" not in line and "-h" not in line.split():
                    cmd = line[line.index("./rdcu"):]
                    after = txt[txt.index(line) + len(line):]
                    eq = [l for l in after.split("\n") if l.startswith("u0'=") or l.startswith("u1'=")][:2]
                    prog = re.findall(r"^([0-9.]+)\t([0-9.e+-]+)\t([0-9.e+-]+)\t([0-9.]+)\t(\d+)\t",
                                      after.split("runtime")[0], flags=re.M)
                    rt = re.search(r"runtime: ([0-9.]+)", after)
                    if prog:
                        last = prog[-1]
                        res["runs"].append({"cell": i, "cmd": cmd, "accepted": int(last[4]),
                                            "final_h": float(last[3]),
                                            "runtime_s": float(rt.group(1)) if rt else None})
                    if eq and "-N 256,256" in cmd and res["equations_2d"] is None:
                        res["equations_2d"] = eq
    json.dump(res, open(OUT, "w"), indent=1)
    print(json.dumps(res, indent=1))


if __name__ == "__main__":
    main()
