"""Writes square8k_clusters.dat: the colour clusters of the square8K mesh
(square8k_mesh.npz) as the oracle's restatement of the reference's greedy
vertex colouring produces them (examples/example2_async_iter.cpp:572-646),
stored in the reference's on-disk form -- the raw uint32 packed CSR of
upload_2d_csr_from (:99-127), written as save_to_binary
(src/Utils/mesh_reader.h:321-329) writes an array.

    python tests/golden/make_square8k_clusters.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402


def main():
    d = np.load(os.path.join(HERE, "square8k_mesh.npz"))
    m = O.Mesh(d["x"], d["faces"])
    csr = np.ascontiguousarray(m.csr(O.CSR_CLUSTERS), np.uint32)
    csr.tofile(os.path.join(HERE, "square8k_clusters.dat"))
    print("clusters", int(csr[0]) - 1, "words", len(csr))


if __name__ == "__main__":
    main()
