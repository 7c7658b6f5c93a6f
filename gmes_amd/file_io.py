"""Probe files and optional dumps (gmes/file_io.py:9-69).

A ``Probe`` here writes the values the engine recorded on the device;
the text format is the reference's (``# location=``, ``# dt=``, then
``n value`` lines).
"""

from os.path import exists
from pathlib import Path
from sys import stderr

import numpy as np


class Probe(object):
    def __init__(self, idx, field, filename):
        self.idx = tuple(idx)
        self.field = field
        f_name = str(filename)
        if exists(f_name):
            stderr.write("Warning: " + f_name + " already exists.\n")
        try:
            self.f = open(f_name, "w")
        except IOError:
            self.f = None
            print("Warning: Can't open file " + f_name + ".\n")

    def __del__(self):
        self.close()

    def close(self):
        if getattr(self, "f", None) is not None and not self.f.closed:
            self.f.close()

    def write_header(self, p, dt):
        self.f.write("# location=" + str(p) + "\n")
        self.f.write("# dt=" + str(dt) + "\n")

    def write(self, n):
        self.f.write(str(n) + " " + str(self.field[self.idx]) + "\n")

    def write_value(self, n, value):
        """One recorded sample, formatted as the reference's numpy scalar."""
        self.f.write(str(n) + " " + str(np.asarray(value, dtype=self._dtype)[()]) + "\n")
        self.f.flush()


def write_hdf5(data, name, low_index, high_index):
    from tables import open_file
    node_name = Path(name).name
    selection = data[low_index[0]:high_index[0], low_index[1]:high_index[1],
                     low_index[2]:high_index[2]]
    with open_file(name + ".h5", mode="w") as h5file:
        h5file.create_array(h5file.root, node_name, selection)


def snapshot(data, filename, title):
    import matplotlib.pyplot as plt
    plt.imsave(filename + "_" + title + ".png", np.real(data))
