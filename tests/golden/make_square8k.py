"""Writes the square8K fixtures from the reference checkout's data files.

Run where the reference checkout is present (tests read only the fixtures):
    python tests/golden/make_square8k.py

Outputs (data only, no reference source):
  square8k_mesh.npz       positions (float32, 8282x3) and triangle faces
                          (int32, 16154x3) parsed from
                          resources/models/square8K.obj (`v` and `f` records,
                          1-based indices, faces with a repeated vertex dropped
                          as mesh_reader.cpp:172-182 does).
  square8k_published.json the energy histories printed by the reference run
                          (documents/example2_async_convergence.py:78-85) and
                          the cluster count of its schedules
                          (documents/example2_scheduling_result.py:310-313).
"""
import json
import os
import re

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def parse_obj(path):
    xs, faces = [], []
    with open(path) as f:
        for line in f:
            t = line.split()
            if not t:
                continue
            if t[0] == "v":
                xs.append([float(t[1]), float(t[2]), float(t[3])])
            elif t[0] == "f":
                ids = [int(s.split("/")[0]) for s in t[1:]]
                ids = [i - 1 if i > 0 else len(xs) + i for i in ids]
                for k in range(1, len(ids) - 1):      # fan triangulation of polygons
                    tri = [ids[0], ids[k], ids[k + 1]]
                    if len(set(tri)) == 3:
                        faces.append(tri)
    return np.array(xs, np.float32), np.array(faces, np.int32)


def parse_list(text, name):
    m = re.search(name + r"\s*=\s*\[(.*?)\]", text, re.S)
    return [float(v) for v in m.group(1).replace("\n", " ").split(",") if v.strip()]


def main():
    x, f = parse_obj(os.path.join(REF, "resources/models/square8K.obj"))
    np.savez_compressed(os.path.join(HERE, "square8k_mesh.npz"), x=x, faces=f)
    conv = open(os.path.join(REF, "documents/example2_async_convergence.py")).read()
    out = {
        "source": "documents/example2_async_convergence.py:78-85",
        "list_sync": parse_list(conv, "list_sync"),
        "list_async": parse_list(conv, "list_async"),
        "num_clusters": 10,
        "num_clusters_source": "documents/example2_scheduling_result.py:310-313",
    }
    with open(os.path.join(HERE, "square8k_published.json"), "w") as fp:
        json.dump(out, fp, indent=1)
    print("verts", len(x), "faces", len(f), "sync", len(out["list_sync"]), "async", len(out["list_async"]))


if __name__ == "__main__":
    main()
