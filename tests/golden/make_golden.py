"""Write the golden fixtures in tests/golden/ from reference runs.

Runs the reference CPU path (oracle/_ref/refprobe[64], built by
`make -C oracle ref` from the reference sources) at the given SCALEs and stores
only DATA: SHA-256 of the generated tuple stream and of each root's level
array, the roots, per-level directions and frontier sizes, traversed-edge
counts, graph counts.  Also copies a few entries of the reference's generated
MRG skip table (generator/mrg_transitions.c) as vectors.

Usage: python tests/golden/make_golden.py [--ref /root/reference] [SCALE ...]
Needs MPICH (mpirun) and the reference sources; never runs on the GPU box.
"""
import argparse
import hashlib
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REFBIN = os.path.join(REPO, "oracle", "_ref")
MPIRUN = "/opt/conda/bin/mpirun"


def run_probe(scale, roots64, threads, spec2012=False):
    exe = os.path.join(REFBIN, "refprobe64" if roots64 else "refprobe")
    d = tempfile.mkdtemp(prefix="g5gold%d_" % scale, dir="/tmp")
    with open(os.path.join(d, "stderr.txt"), "w") as err:
        subprocess.check_call([MPIRUN, "-np", "1", "-genv", "OMP_NUM_THREADS", str(threads), exe,
                               str(scale), "16", d] + (["spec2012"] if spec2012 else []),
                              stdout=subprocess.DEVNULL, stderr=err)
    return d


def summarize(scale, d, nroots_expected):
    out = {"scale": scale, "edgefactor": 16, "seeds": [2, 3]}
    t = np.fromfile(os.path.join(d, "tuples.bin"), np.uint32).reshape(-1, 3)
    out["num_tuples"] = int(t.shape[0])
    out["tuples_sha256"] = hashlib.sha256(t.tobytes()).hexdigest()
    out["tuples_head"] = t[:8].tolist()
    out["tuples_tail"] = t[-8:].tolist()
    info = dict(l.split() for l in open(os.path.join(d, "info.txt")))
    out.update({k: int(v) for k, v in info.items()})
    roots = [int(x) for x in open(os.path.join(d, "roots.txt")).read().split()]
    assert len(roots) == nroots_expected
    out["roots"] = roots
    nedge = [l.split() for l in open(os.path.join(d, "nedge.txt"))]
    blocks = open(os.path.join(d, "stderr.txt")).read().split("refprobe: root ")[1:]
    per = []
    for i, r in enumerate(roots):
        lv = np.fromfile(os.path.join(d, "levels_%d.bin" % i), np.uint8)
        dirs = re.findall(r"Direction (top-down|bottom-up)\n", blocks[i])
        nqs = [int(x) for x in re.findall(r"NQ (\d+),", blocks[i])]
        hist = np.bincount(lv[lv != 0xFF]).tolist()
        per.append({
            "root": r,
            "valid": int(nedge[i][0]),
            "nedge": int(nedge[i][1]),
            "levels_sha256": hashlib.sha256(lv.tobytes()).hexdigest(),
            "level_histogram": hist,
            "directions": "".join("T" if x == "top-down" else "B" for x in dirs),
            "nq": nqs,
        })
    out["per_root"] = per
    return out


def mrg_table_vectors(ref):
    """A handful of (byte, value) -> 9 entries of the reference skip table."""
    txt = open(os.path.join(ref, "generator", "mrg_transitions.c")).read()
    body = txt[txt.index("mrg_skip_matrices[][256] = {"):]
    parts = re.split(r"/\* Byte (\d+) \*/", body)[1:]
    rows = {}
    for bi in range(0, len(parts), 2):
        b = int(parts[bi])
        ents = re.findall(r"\{(\d+), (\d+), (\d+), (\d+), (\d+), (\d+), (\d+), (\d+), (\d+)\}",
                          parts[bi + 1])
        rows[b] = [[int(x) for x in e] for e in ents]
    want = [(0, 0), (0, 1), (0, 5), (0, 255), (1, 1), (1, 77), (7, 200), (8, 1), (8, 2),
            (8, 255), (9, 3), (11, 128), (15, 255), (16, 2), (16, 50), (17, 7), (23, 255)]
    return [{"byte": b, "value": v, "entry": rows[b][v]} for b, v in want]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("scales", nargs="*", type=int, default=[16, 20])
    ap.add_argument("--roots64", type=int, nargs="*", default=[16],
                    help="SCALEs that also get a 64-root fixture (REAL_BENCHMARK root count)")
    ap.add_argument("--only64", action="store_true", help="write only the 64-root fixtures")
    ap.add_argument("--spec2012", type=int, nargs="*", default=[],
                    help="SCALEs for spec-2012 fixtures (initiator 5500/100, binary-tree initial edges)")
    a = ap.parse_args()
    if not os.path.exists(os.path.join(REFBIN, "refprobe")):
        sys.exit("build the reference probe first: make -C oracle ref")
    json.dump({"source": "generator/mrg_transitions.c", "vectors": mrg_table_vectors(a.ref)},
              open(os.path.join(HERE, "mrg_table.json"), "w"), indent=1)
    for s in a.scales:
        for roots64 in (False, True):
            if roots64 and s not in a.roots64:
                continue
            if not roots64 and a.only64:
                continue
            d = run_probe(s, roots64, a.threads)
            res = summarize(s, d, 64 if roots64 else 16)
            name = "ref_s%d_%droots.json" % (s, 64 if roots64 else 16)
            json.dump(res, open(os.path.join(HERE, name), "w"), indent=1)
            subprocess.call(["rm", "-rf", d])
            print("wrote", name)
    for s in a.spec2012:
        d = run_probe(s, False, a.threads, spec2012=True)
        res = summarize(s, d, 16)
        res["generator"] = "spec2012: initiator 5500/100 per 10000, 2^SCALE-1 binary-tree initial edges first"
        name = "ref_spec2012_s%d_16roots.json" % s
        json.dump(res, open(os.path.join(HERE, name), "w"), indent=1)
        subprocess.call(["rm", "-rf", d])
        print("wrote", name)


if __name__ == "__main__":
    main()
