"""Leaf-block target format: nb^3 cells per cubic block, implicit cell centres.

Block b has lower corner origin[b] and cell width dx[b]; cell (i, j, k) has centre
origin + (idx + 1/2) dx and flat index b*nb^3 + (i*nb + j)*nb + k -- the order of
``np.meshgrid(x, y, z, indexing='ij')`` flattened with ``ravel('C')`` that the reference's
per-block bridge uses (.deprecated/bridge_multiples.py:566-571) for FLASH's
get_1blk_cell_coords (src/amuse/torch_amuse_flash/interface.F90:418-437).
The device kernels rebuild the same centres bit for bit (multiply, then add).
"""
import numpy as np


def block_centres(origin, dx, nb=8):
    """(x, y, z, dx_per_cell) of all cells, numpy float64."""
    origin = np.asarray(origin, dtype=np.float64).reshape(-1, 3)
    dx = np.asarray(dx, dtype=np.float64).reshape(-1)
    idx = np.arange(nb, dtype=np.float64) + 0.5
    ci = np.broadcast_to(idx[:, None, None], (nb, nb, nb)).ravel()
    cj = np.broadcast_to(idx[None, :, None], (nb, nb, nb)).ravel()
    ck = np.broadcast_to(idx[None, None, :], (nb, nb, nb)).ravel()
    x = (origin[:, 0:1] + ci[None, :] * dx[:, None]).ravel()
    y = (origin[:, 1:2] + cj[None, :] * dx[:, None]).ravel()
    z = (origin[:, 2:3] + ck[None, :] * dx[:, None]).ravel()
    return x, y, z, np.repeat(dx, nb ** 3)
