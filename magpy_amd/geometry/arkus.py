"""Arkus cluster geometries (magpy/geometry/arkus.py).

ARKUS[n_particles][configuration_id] is an (n_particles, 3) array of
particle centres at unit contact distance: the minimum-energy sticky-sphere
clusters of N. Arkus, V. N. Manoharan and M. P. Brenner, Phys. Rev. Lett.
103, 118303 (2009), for 1-8 particles (9 has no entries), in the
reference's ordering and orientation.  The numbers live in arkus.json
(written by tests/golden/make_arkus_table.py).
"""
import json
import os

import numpy as np

with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "arkus.json")) as _f:
    ARKUS = {int(n): {int(c): np.array(xyz, dtype=np.float64) for c, xyz in confs.items()}
             for n, confs in json.load(_f).items()}
