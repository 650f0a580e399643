#!/usr/bin/env python3
"""Regenerate the golden fixtures of tests/golden/ (run where the reference
tree is mounted; GPU test runs only read the fixtures it wrote).

Two instances, both with 10 metal layers because the reference reads only
10-layer grids (src/input.hpp:292):
  syn10    synthetic instance of workload configuration 7 (64x64x10, 1500
           nets, 30% of pins with 2-3 access points), seed 1;
  mempool  a 96x96 window of the reference's own input fixture
           baseline/InstantGR2/benchmarks/mempool_tile_rank.{cap,net}: the
           nets whose access points all lie inside, re-based to the window.

For each, the CAP/NET pair and the output of oracle/_ref/ref_harness (the
reference's own host code: pin selection, FLUTE trees, subnet split, hpwl
order and CPU batch lists, one JSON object per line) are written gzipped.

`--flt` adds {name}.flt.jsonl.gz from the committed CAP/NET pairs: a
congestion view (first line, {"view": {"xsum", "ysum"}}: a fixed pattern of
congested cells and its exclusive prefix sums along x and y) and the FLT
DAG that net::generate_detours_reconstruct (src/preprocess.hpp:769-1436)
builds from it for every subnet tree of two or more nodes.
"""
import gzip
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
sys.path.insert(0, ROOT)


def write_gz(path, text):
    with gzip.open(path, "wt") as f:
        f.write(text)


def crop_mempool(x0, y0, side, max_nets):
    base = os.path.join(REF, "baseline/InstantGR2/benchmarks/mempool_tile_rank")
    with open(base + ".cap") as f:
        lines = f.read().split("\n")
    L, X, Y = map(int, lines[0].split())
    costs = lines[1].split()
    xlen = lines[2].split()
    ylen = lines[3].split()
    out = [f"{L} {side} {side}", " ".join(costs),
           " ".join(xlen[x0:x0 + side - 1]), " ".join(ylen[y0:y0 + side - 1])]
    row = 4
    for _ in range(L):
        out.append(lines[row])
        row += 1
        for y in range(Y):
            if y0 <= y < y0 + side:
                vals = lines[row].split()
                out.append(" ".join(vals[x0:x0 + side]))
            row += 1
    cap = "\n".join(out) + "\n"
    with open(base + ".net") as f:
        text = f.read()
    nets = []
    for m in re.finditer(r"(\S+)\n\(\n(.*?)\n\)", text, re.S):
        name, body = m.group(1), m.group(2)
        pins = []
        ok = True
        for line in body.split("\n"):
            pts = [tuple(map(int, t)) for t in re.findall(r"\((\d+), (\d+), (\d+)\)", line)]
            if not pts or any(not (x0 <= x < x0 + side and y0 <= y < y0 + side) for _, x, y in pts):
                ok = False
                break
            pins.append("[" + ", ".join(f"({l}, {x - x0}, {y - y0})" for l, x, y in pts) + "]")
        if ok and len(pins) >= 1:
            nets.append(f"{name}\n(\n" + "\n".join(pins) + "\n)\n")
        if len(nets) >= max_nets:
            break
    return cap, "".join(nets)


def run_harness(cap_text, net_text, name):
    tmp = f"/tmp/golden_{name}"
    with open(tmp + ".cap", "w") as f:
        f.write(cap_text)
    with open(tmp + ".net", "w") as f:
        f.write(net_text)
    res = subprocess.run([os.path.join(ROOT, "oracle/_ref/ref_harness"), tmp + ".cap", tmp + ".net",
                          os.path.join(REF, "resources/flute")], capture_output=True, text=True, check=True)
    return res.stdout


def flt_view(X, Y):
    """Congested cells on a fixed pattern (worst overflow 1 or 2) and the
    exclusive prefix sums along x and along y, float32, x*Y+y order."""
    import numpy as np
    x, y = np.meshgrid(np.arange(X), np.arange(Y), indexing="ij")
    worst = np.where((x * 131 + y * 71) % 17 == 0, 1.0 + (x + y) % 2, 0.0).astype(np.float32)
    xsum = np.zeros((X, Y), np.float32)
    ysum = np.zeros((X, Y), np.float32)
    for i in range(1, X):
        xsum[i] = xsum[i - 1] + worst[i - 1]
    for j in range(1, Y):
        ysum[:, j] = ysum[:, j - 1] + worst[:, j - 1]
    return xsum.reshape(-1), ysum.reshape(-1)


def make_flt(name):
    import gzip as gz
    import json
    with gz.open(os.path.join(HERE, f"{name}.cap.gz"), "rt") as f:
        cap = f.read()
    with gz.open(os.path.join(HERE, f"{name}.net.gz"), "rt") as f:
        net = f.read()
    _, X, Y = map(int, cap.split("\n")[0].split())
    xsum, ysum = flt_view(X, Y)
    tmp = f"/tmp/golden_{name}"
    with open(tmp + ".cap", "w") as f:
        f.write(cap)
    with open(tmp + ".net", "w") as f:
        f.write(net)
    with open(tmp + ".view", "wb") as f:
        f.write(xsum.tobytes() + ysum.tobytes())
    res = subprocess.run([os.path.join(ROOT, "oracle/_ref/ref_harness"), tmp + ".cap", tmp + ".net",
                          os.path.join(REF, "resources/flute"), "1", tmp + ".view"],
                         capture_output=True, text=True, check=True)
    dags = [line for line in res.stdout.split("\n") if line.startswith('{"dag"')]
    head = json.dumps({"view": {"xsum": [float(v) for v in xsum], "ysum": [float(v) for v in ysum]}})
    write_gz(os.path.join(HERE, f"{name}.flt.jsonl.gz"), head + "\n" + "\n".join(dags) + "\n")
    print(name, "FLT DAGs", len(dags))


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    if "--flt" in sys.argv:
        for name in ("syn10", "mempool"):
            make_flt(name)
        return
    import grotor_amd as ga
    d = ga.Design.generate(7, seed=1)
    d.write("/tmp/golden_syn10_src.cap", "/tmp/golden_syn10_src.net")
    instances = {
        "syn10": (open("/tmp/golden_syn10_src.cap").read(), open("/tmp/golden_syn10_src.net").read()),
        "mempool": crop_mempool(160, 200, 96, 2000),
    }
    for name, (cap, net) in instances.items():
        dump = run_harness(cap, net, name)
        write_gz(os.path.join(HERE, f"{name}.cap.gz"), cap)
        write_gz(os.path.join(HERE, f"{name}.net.gz"), net)
        write_gz(os.path.join(HERE, f"{name}.ref.jsonl.gz"), dump)
        print(name, "nets", net.count("\n(\n"), "dump lines", dump.count("\n"))


if __name__ == "__main__":
    main()
