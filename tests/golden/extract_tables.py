"""Writes tests/golden/dc_tables.npz: the dual-contouring lookup tables as
the reference defines them, for checking the oracle's and the product's
tables.

* corner_manifold[256]: the table printed in
  libfive/src/render/brep/dc/dc_tree3.cpp:146-153, read as data.
* marching_v[256, 4, 12, 2], marching_e[8, 8], marching_p[256, 24]: the
  marching table of libfive/src/render/brep/dc/marching.cpp, expanded here
  in Python from the seed cases listed in loadCases<3> (read as data from
  that file): quarter turns about X, Y, Z applied until nothing changes,
  then directed-edge ids in first-seen order (marching.cpp:420-527).

Needs the reference checkout (default /root/reference, or the first
argument); run it where that exists:  python tests/golden/extract_tables.py
"""
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def corner_manifold(ref):
    src = open(os.path.join(ref, "libfive/src/render/brep/dc/dc_tree3.cpp")).read()
    body = src[src.index("corner_table[] ="):]
    body = body[body.index("{") + 1:body.index("}")]
    bits = [int(t) for t in re.findall(r"[01]", body)]
    assert len(bits) == 256, len(bits)
    return np.array(bits, np.uint8)


def seed_cases(ref):
    src = open(os.path.join(ref, "libfive/src/render/brep/dc/marching.cpp")).read()
    start = src.index("void loadCases<3>")
    body = src[start:src.index("\n}\n", start)]
    cases = []
    for line in body.splitlines():
        line = line.strip()
        m = re.match(r"begin\(\{([\d, ]+)\}\);", line)
        if m:
            cases.append(([int(x) for x in m.group(1).split(",")], [[]]))
            continue
        m = re.match(r"push\(\{(\d+), ?(\d+)\}\);", line)
        if m:
            cases[-1][1][-1].append((int(m.group(1)), int(m.group(2))))
            continue
        if line == "next();":
            cases[-1][1].append([])
    return cases


def rotate_vertex(v, axis):
    c = [1 if v & (1 << i) else -1 for i in range(3)]
    if axis == 0:      # +90 degrees about X: (x, y, z) -> (x, -z, y)
        r = (c[0], -c[2], c[1])
    elif axis == 1:    # about Y: (x, y, z) -> (z, y, -x)
        r = (c[2], c[1], -c[0])
    else:              # about Z: (x, y, z) -> (-y, x, z)
        r = (-c[1], c[0], c[2])
    return sum(1 << i for i in range(3) if r[i] > 0)


def marching(cases):
    v = -np.ones((256, 4, 12, 2), np.int8)
    for verts, patches in cases:
        mask = sum(1 << i for i in verts)
        for p, edges in enumerate(patches):
            for e, (a, b) in enumerate(edges):
                v[mask, p, e] = (a, b)
    changed = [bool(v[i, 0, 0, 0] != -1) for i in range(256)]
    any_changed = True
    while any_changed:
        any_changed = False
        for i in range(256):
            if not changed[i]:
                continue
            changed[i] = False
            for axis in range(3):
                j = sum(1 << rotate_vertex(k, axis) for k in range(8) if i & (1 << k))
                if v[j, 0, 0, 0] != -1:
                    continue
                changed[j] = True
                any_changed = True
                for p in range(4):
                    if v[i, p, 0, 0] == -1:
                        break
                    for e in range(12):
                        if v[i, p, e, 0] == -1:
                            break
                        v[j, p, e] = (rotate_vertex(int(v[i, p, e, 0]), axis),
                                      rotate_vertex(int(v[i, p, e, 1]), axis))
    e_ids = -np.ones((8, 8), np.int8)
    p_ids = -np.ones((256, 24), np.int8)
    nxt = 0
    for i in range(256):
        for p in range(4):
            if v[i, p, 0, 0] == -1:
                break
            for e in range(12):
                if v[i, p, e, 0] == -1:
                    break
                a, b = int(v[i, p, e, 0]), int(v[i, p, e, 1])
                if e_ids[a, b] == -1:
                    e_ids[a, b] = nxt
                    nxt += 1
                p_ids[i, e_ids[a, b]] = p
    assert nxt == 24
    return v, e_ids, p_ids


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    cm = corner_manifold(ref)
    v, e, p = marching(seed_cases(ref))
    out = os.path.join(HERE, "dc_tables.npz")
    np.savez_compressed(out, corner_manifold=cm, marching_v=v, marching_e=e,
                        marching_p=p)
    print("wrote", out)


if __name__ == "__main__":
    main()
