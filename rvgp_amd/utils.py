"""Drop-in for RVGP/utils.py: Wavefront .obj mesh loading.

``load_mesh(data, folder)`` reads ``<folder>/<data>.obj`` exactly as
RVGP/utils.py:7-47 does: lines starting with '#' are skipped, 'v x y z'
lines give vertices (float64), 'f i j k' lines give faces (1-based integers,
returned 0-based); every other record is ignored.  Like the reference, a
blank line raises IndexError and a face with '/'-separated texture or
normal indices raises ValueError.  The reference's default folder is its
own examples/data directory; this package ships no meshes, so the default
folder is $RVGP_MESH_DIR (or the current directory).
"""
import os

import numpy as np


def load_mesh(data='bunny', folder=None):
    """(vertices (n, 3) float64, faces (m, 3) int, 0-based)."""
    if folder is None:
        folder = os.environ.get("RVGP_MESH_DIR", os.getcwd())
    path = '{}.obj'.format(os.path.join(folder, data))
    vertices, faces = [], []
    with open(path, 'r') as fh:
        for line in fh:
            if line.startswith('#'):
                continue
            words = line.split()
            if words[0] == 'v':
                vertices.append([float(words[1]), float(words[2]), float(words[3])])
            elif words[0] == 'f':
                faces.append([int(words[1]), int(words[2]), int(words[3])])
    return np.array(vertices), np.array(faces) - 1
