"""Loaders for the golden fixtures of tests/golden/ (made by make_golden.py
from the reference's own host code)."""
import functools
import gzip
import json
import os

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ("syn10", "mempool")


def path(name, ext):
    return os.path.join(HERE, f"{name}.{ext}.gz")


def unpacked(name, ext):
    """Decompress a fixture into /tmp and return the path."""
    out = f"/tmp/grotor_golden_{name}.{ext}"
    if not os.path.exists(out):
        with gzip.open(path(name, ext), "rt") as f, open(out, "w") as g:
            g.write(f.read())
    return out


@functools.lru_cache(maxsize=None)
def dump(name):
    d = {"nets": [], "subs": []}
    with gzip.open(path(name, "ref.jsonl"), "rt") as f:
        for line in f:
            o = json.loads(line)
            if "grid" in o:
                d["grid"] = o["grid"]
            elif "net" in o:
                d["nets"].append(o)
            elif "sub" in o:
                d["subs"].append(o)
            elif "order" in o:
                d["order"] = o["order"]
            elif "batches" in o:
                d["batches"] = o["batches"]
    return d


@functools.lru_cache(maxsize=None)
def grid(name):
    """Routing grid of a golden CAP file, in the shape Context() takes."""
    with gzip.open(path(name, "cap"), "rt") as f:
        lines = f.read().split("\n")
    L, X, Y = map(int, lines[0].split())
    head = list(map(float, lines[1].split()))
    xlen = np.array(lines[2].split(), np.int32)
    ylen = np.array(lines[3].split(), np.int32)
    cap = np.zeros((L, X, Y), np.float32)
    dirs = []
    row = 4
    for l in range(L):
        dirs.append(int(lines[row].split()[1]))
        row += 1
        for y in range(Y):
            cap[l, :, y] = np.array(lines[row].split(), np.float32)
            row += 1
    return dict(L=L - 1, X=X, Y=Y, DIR=dirs[1], unit_length_wire_cost=head[0], unit_via_cost=head[1],
                unit_length_short_costs=np.array(head[3:3 + L - 1], np.float64), x_edge_len=xlen, y_edge_len=ylen,
                capacity=np.ascontiguousarray(cap[1:]).reshape(-1))


def preorder(nodes, adj):
    """DFS preorder from node 0 (src/routing/lshape.cuh:86-102)."""
    out_nodes, out_par = [], []
    stack = [(0, -1, 0, iter(adj[0]) if adj else iter(()))]
    out_nodes.append(nodes[0])
    out_par.append(-1)
    idx_of = {0: 0}
    while stack:
        x, par, idx, it = stack[-1]
        nxt = next(it, None)
        if nxt is None:
            stack.pop()
            continue
        if nxt == par:
            continue
        i = len(out_nodes)
        out_nodes.append(nodes[nxt])
        out_par.append(idx)
        idx_of[nxt] = i
        stack.append((nxt, x, i, iter(adj[nxt])))
    return out_nodes, out_par


@functools.lru_cache(maxsize=None)
def topology(name):
    """Stage arrays in the reference's priority order, from its own trees."""
    d = dump(name)
    g = grid(name)
    subs = d["subs"]
    order = d["order"]
    node_off, nodes, parents = [0], [], []
    pt_off, pts, hs_off, hsegs, vs_off, vsegs = [0], [], [0], [], [0], []
    for sid in order:
        s = subs[sid]
        nd, pr = preorder(s["nodes"], s["adj"])
        nodes += nd
        parents += pr
        node_off.append(len(nodes))
        pts += s["points"]
        pt_off.append(len(pts))
        hsegs += s["hsegs"]
        hs_off.append(len(hsegs) // 2)
        vsegs += s["vsegs"]
        vs_off.append(len(vsegs) // 2)
    pin_acc = np.concatenate([[0], np.cumsum([len(s["pins"]) for s in subs])]).astype(np.int32)
    t = dict(g)
    t.update(nets=len(order), subnets=len(subs), order=np.array(order, np.int32),
             node_off=np.array(node_off, np.int32), nodes=np.array(nodes, np.int32),
             parents=np.array(parents, np.int32), pt_off=np.array(pt_off, np.int32), pts=np.array(pts, np.int32),
             hs_off=np.array(hs_off, np.int32), hsegs=np.array(hsegs, np.int32), vs_off=np.array(vs_off, np.int32),
             vsegs=np.array(vsegs, np.int32), pin_acc_num=pin_acc)
    return t


def reference_colors(name):
    """Batch index per priority slot from the reference's batch lists."""
    d = dump(name)
    slot_of = {sid: k for k, sid in enumerate(d["order"])}
    color = np.full(len(d["order"]), -1, np.int32)
    for b, lst in enumerate(d["batches"]):
        for sid in lst:
            color[slot_of[sid]] = b
    return color, len(d["batches"])


def write_out(name, routes, pin_acc_num, out_path):
    """.out text for the fixture's nets (src/output.hpp:79-112)."""
    d = dump(name)
    X, Y = d["grid"]["X"], d["grid"]["Y"]
    XY = X * Y
    lines = []

    def seg(a, b):
        if a > b:
            a, b = b, a
        lines.append(f"{a // Y % X} {a % Y} {a // XY} {b // Y % X} {b % Y} {b // XY}")

    for n in d["nets"]:
        lines.append(n["name"])
        lines.append("(")
        for p in n["pins"]:
            if p < XY:
                seg(p, p + XY)
        ex = n["extra"]
        for i in range(0, len(ex) - 1, 2):
            seg(ex[i], ex[i + 1])
        for sid in n["subnets"]:
            base = 12 * int(pin_acc_num[sid])
            r = routes[base:]
            for j in range(1, int(r[0]), 2):
                seg(int(r[j]) + XY, int(r[j + 1]) + XY)
        lines.append(")")
    with open(out_path, "w") as f:
        f.write("\n".join(lines) + "\n")
