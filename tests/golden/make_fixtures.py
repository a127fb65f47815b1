"""Write the small data fixtures under tests/golden/ from the reference's data files.

Reads ``circle1.mat`` (a FEM mesh of the unit disc the reference ships: K,
lumped M, nodes, elements; 468 nodes) with scipy.io.loadmat, which only
parses MAT-v5 data, and stores the arrays as ``circle1.npz``.  Run in the
container that has /root/reference:

    python tests/golden/make_fixtures.py
"""

import os

import numpy as np
import scipy.io as sio
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/circle1.mat"


def main():
    d = sio.loadmat(SRC)
    K = sp.coo_matrix(d["K"])
    M = sp.coo_matrix(d["M"])
    np.savez_compressed(
        os.path.join(HERE, "circle1.npz"),
        nodes=np.asarray(d["nodes"], dtype=np.float64),
        elements=np.asarray(d["elements"], dtype=np.int32),      # 1-based, as stored
        K_row=K.row.astype(np.int32), K_col=K.col.astype(np.int32), K_val=K.data,
        M_diag=M.diagonal().astype(np.float64),
        boundary_nodes=np.asarray(d["boundary_nodes"], dtype=np.int32).ravel(),
    )
    print("wrote circle1.npz")


if __name__ == "__main__":
    main()
