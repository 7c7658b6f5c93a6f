"""Probe files and image dumps.

The engine samples probed cells on the device every half step into a ring
buffer; the host drains it after each run of steps and appends the
samples to one text file per probe.  The file format is the reference's
(gmes/file_io.py:9-48), since scripts read it: a ``# location=`` line, a
``# dt=`` line, then one ``n value`` line per sample with the value printed
as a numpy scalar of the field's dtype.
"""

import os
import sys

import numpy as np


class ProbeFile(object):
    """The text file of one probe: the header at creation, then samples
    appended in batches as they are read back from the device."""

    def __init__(self, path, location, dt, dtype):
        path = str(path)
        self.path = path
        self.dtype = np.dtype(dtype)
        self.pid = None      # the engine's probe id, set by the simulation
        # the reference's warnings, word for word (gmes/file_io.py:12-20)
        if os.path.exists(path):
            sys.stderr.write("Warning: " + path + " already exists.\n")
        try:
            self._fh = open(path, "w")
        except IOError:
            self._fh = None
            print("Warning: Can't open file " + path + ".\n")
            return
        self._fh.write("# location=%s\n# dt=%s\n" % (location, dt))
        self._fh.flush()

    def append(self, steps, values):
        """Append samples: steps[i] is the half-step index n of values[i]."""
        if self._fh is None or len(values) == 0:
            return
        vals = np.asarray(values).astype(self.dtype, copy=False)
        self._fh.write("".join("%s %s\n" % (n, v) for n, v in zip(steps, vals)))
        self._fh.flush()

    def close(self):
        if self._fh is not None and not self._fh.closed:
            self._fh.close()

    def __del__(self):
        self.close()


def save_png(data, filename, title):
    """A plane of a field as an image (needs matplotlib)."""
    import matplotlib.pyplot as plt
    plt.imsave(filename + "_" + title + ".png", np.real(data))
