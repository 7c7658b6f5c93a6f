"""The Cartesian grid: sizes, staggered storage shapes and index maps.

Drop-in counterpart of the reference's ``gmes.geometry`` module
(gmes/geometry.py:16-812).  The engine runs one simulation per GPU, so
the communicator is always the serial periodic one of the reference
(``AuxiCartComm``, gmes/geometry.py:16-101); ``parallel=True`` is
accepted and runs serially, as the reference does when mpi4py is absent
(gmes/geometry.py:155-165).

The index <-> space maps keep the reference's floating-point expressions
(e.g. ``(i + 0.5) * dx - half_size``) so that every staggered coordinate,
and therefore every material decision, is bit-identical.  ``axis_coords``
returns the same numbers for a whole axis at once.
"""

import numpy as np
from numpy import array, empty, inf

from . import constant as const
from .pygeom import (  # noqa: F401  (re-exported like the reference)
    Block,
    Compound,
    Cone,
    Cylinder,
    DefaultMedium,
    Ellipsoid,
    GeomBox,
    GeomBoxTree,
    GeometricObject,
    Material,
    Shell,
    Sphere,
)

# Half-cell offset of each component's storage index along x, y, z
# (gmes/geometry.py:360-732): Ex (+.5,0,0) ... Hz (-.5,-.5,0).
STAGGER = {
    const.Ex: (0.5, 0.0, 0.0),
    const.Ey: (0.0, 0.5, 0.0),
    const.Ez: (0.0, 0.0, 0.5),
    const.Hx: (0.0, -0.5, -0.5),
    const.Hy: (-0.5, 0.0, -0.5),
    const.Hz: (-0.5, -0.5, 0.0),
}

# Storage index used on a collapsed (size-1) axis by spc_to_exact_<c>_idx
# (gmes/geometry.py:391-400, 594-602, 645-652, 696-703).
COLLAPSED_INDEX = {
    const.Ex: (0, 0, 0),
    const.Ey: (0, 0, 0),
    const.Ez: (0, 0, 0),
    const.Hx: (0, 1, 1),
    const.Hy: (1, 0, 1),
    const.Hz: (1, 1, 0),
}


class AuxiCartComm(object):
    """Serial stand-in of a periodic MPI Cartesian communicator
    (gmes/geometry.py:16-101): every shift points back to this rank."""

    def __init__(self, dims=(1, 1, 1), periods=None, reorder=(0, 0, 0)):
        self.rank = 0
        self.dim = 3
        cyclic = tuple(map(int, periods)) if periods else (0, 0, 0)
        self.topo = ((1, 1, 1), cyclic, (0, 0, 0))

    def Get_cart_rank(self, coords):
        return 0

    def Get_coords(self, rank):
        return 0, 0, 0

    def Get_dim(self):
        return self.dim

    def Get_topo(self):
        return self.topo

    def Get_size(self):
        return 1

    def Shift(self, direction, disp):
        return (0, 0) if self.topo[1][direction] else (-1, -1)

    def sendrecv(self, sendbuf, dest=0, sendtag=0, recvbuf=None, source=0, recvtag=0,
                 status=None):
        return None if dest == -1 or source == -1 else sendbuf

    def reduce(self, value, root=0, op=None):
        return value

    def bcast(self, obj=None, root=0):
        return obj

    def allgather(self, obj=None):
        return [obj]


class Cartesian(object):
    """The calculation space (gmes/geometry.py:104-175).

    Attributes follow the reference: half_size, res, dr, dt,
    whole_field_size, my_id, numprocs, cart_comm, my_cart_idx,
    general_field_size, my_field_size.
    """

    def __init__(self, size, resolution=15, parallel=False):
        try:
            if len(resolution) == 3:
                self.res = array(resolution, np.double)
        except TypeError:
            self.res = array((resolution,) * 3, np.double)

        self.dr = array(1 / self.res, np.double)
        self.half_size = 0.5 * array(size, np.double)
        for i, v in enumerate(self.dr):
            if self.half_size[i] == 0:
                self.half_size[i] = 0.5 * v

        self.whole_field_size = array((2 * self.half_size * self.res).round(), np.intp)
        self.my_id = 0
        self.numprocs = 1
        self.cart_comm = AuxiCartComm((1, 1, 1), (1, 1, 1))
        self.my_cart_idx = self.cart_comm.Get_topo()[2]
        self.general_field_size = self.whole_field_size // self.cart_comm.Get_topo()[0]
        self.my_field_size = self.get_my_field_size()
        self.dt = None

    def bcast(self, obj=None, root=None):
        return obj

    def get_my_field_size(self):
        field_size = empty(3, np.intp)
        dims = self.cart_comm.Get_topo()[0]
        for i in range(3):
            if self.my_cart_idx[i] == dims[i] - 1:
                field_size[i] = self.whole_field_size[i] - (
                    self.my_cart_idx[i] * self.general_field_size[i])
            else:
                field_size[i] = self.general_field_size[i]
        return field_size

    def find_best_deploy(self):
        return (1, 1, 1)

    # ---- storage -------------------------------------------------------

    def storage_shape(self, comp, field_compnt):
        """Shape of the array of ``comp`` (gmes/geometry.py:276-358)."""
        if comp not in field_compnt:
            return (1, 1, 1)
        n = [int(v) for v in self.my_field_size]
        shape = [v + 1 for v in n]
        own = comp.index % 3
        shape[own] = n[own]
        return tuple(shape)

    def _get_storage(self, comp, field_compnt, cmplx, dtype=None):
        if dtype is None:
            dtype = complex if cmplx else np.double
        return np.zeros(self.storage_shape(comp, field_compnt), dtype)

    def get_ex_storage(self, field_compnt, cmplx=False):
        return self._get_storage(const.Ex, field_compnt, cmplx)

    def get_ey_storage(self, field_compnt, cmplx=False):
        return self._get_storage(const.Ey, field_compnt, cmplx)

    def get_ez_storage(self, field_compnt, cmplx=False):
        return self._get_storage(const.Ez, field_compnt, cmplx)

    def get_hx_storage(self, field_compnt, cmplx=False):
        return self._get_storage(const.Hx, field_compnt, cmplx)

    def get_hy_storage(self, field_compnt, cmplx=False):
        return self._get_storage(const.Hy, field_compnt, cmplx)

    def get_hz_storage(self, field_compnt, cmplx=False):
        return self._get_storage(const.Hz, field_compnt, cmplx)

    # ---- index <-> space -----------------------------------------------

    def index_to_space(self, comp, i, j, k):
        """Space coordinate of a storage index of ``comp``."""
        idx = array((i, j, k), np.intp)
        global_idx = idx + self.general_field_size * self.my_cart_idx
        off = STAGGER[comp]
        out = []
        for a in range(3):
            if off[a] == 0:
                out.append(global_idx[a] * self.dr[a] - self.half_size[a])
            else:
                out.append((global_idx[a] + off[a]) * self.dr[a] - self.half_size[a])
        return tuple(out)

    def axis_coords(self, comp, axis, n):
        """Coordinates of storage indices 0..n-1 of ``comp`` along ``axis``
        (the same expression as ``index_to_space``, vectorized)."""
        idx = np.arange(n, dtype=np.intp) + self.general_field_size[axis] * self.my_cart_idx[axis]
        off = STAGGER[comp][axis]
        if off == 0:
            return idx * self.dr[axis] - self.half_size[axis]
        return (idx + off) * self.dr[axis] - self.half_size[axis]

    def spc_to_exact_idx(self, comp, x, y, z):
        """Exact (fractional) storage index of a point
        (gmes/geometry.py:376-400 and the analogous methods)."""
        coords = array((x, y, z), np.double)
        off = STAGGER[comp]
        global_idx = empty(3, np.double)
        for a in range(3):
            g = (coords[a] + self.half_size[a]) / self.dr[a]
            global_idx[a] = g - off[a] if off[a] != 0 else g
        if comp in const.E_COMPONENTS:
            idx = empty(3, np.double)
            for a in range(3):
                if self.whole_field_size[a] == 1:
                    idx[a] = 0
                else:
                    idx[a] = global_idx[a] - (self.my_cart_idx[a] * self.general_field_size[a])
        else:
            idx = global_idx - self.my_cart_idx * self.general_field_size
            col = COLLAPSED_INDEX[comp]
            for a in range(3):
                if self.whole_field_size[a] == 1:
                    idx[a] = col[a]
        return tuple(idx)

    def space_to_index(self, comp, x, y, z):
        """Nearest storage index of ``comp`` (may be out of range)."""
        exact_idx = self.spc_to_exact_idx(comp, x, y, z)
        return tuple(np.floor(array(exact_idx) + 0.5).astype(np.intp))

    def ex_index_to_space(self, i, j, k):
        return self.index_to_space(const.Ex, i, j, k)

    def ey_index_to_space(self, i, j, k):
        return self.index_to_space(const.Ey, i, j, k)

    def ez_index_to_space(self, i, j, k):
        return self.index_to_space(const.Ez, i, j, k)

    def hx_index_to_space(self, i, j, k):
        return self.index_to_space(const.Hx, i, j, k)

    def hy_index_to_space(self, i, j, k):
        return self.index_to_space(const.Hy, i, j, k)

    def hz_index_to_space(self, i, j, k):
        return self.index_to_space(const.Hz, i, j, k)

    def spc_to_exact_ex_idx(self, x, y, z):
        return self.spc_to_exact_idx(const.Ex, x, y, z)

    def spc_to_exact_ey_idx(self, x, y, z):
        return self.spc_to_exact_idx(const.Ey, x, y, z)

    def spc_to_exact_ez_idx(self, x, y, z):
        return self.spc_to_exact_idx(const.Ez, x, y, z)

    def spc_to_exact_hx_idx(self, x, y, z):
        return self.spc_to_exact_idx(const.Hx, x, y, z)

    def spc_to_exact_hy_idx(self, x, y, z):
        return self.spc_to_exact_idx(const.Hy, x, y, z)

    def spc_to_exact_hz_idx(self, x, y, z):
        return self.spc_to_exact_idx(const.Hz, x, y, z)

    def space_to_ex_index(self, x, y, z):
        return self.space_to_index(const.Ex, x, y, z)

    def space_to_ey_index(self, x, y, z):
        return self.space_to_index(const.Ey, x, y, z)

    def space_to_ez_index(self, x, y, z):
        return self.space_to_index(const.Ez, x, y, z)

    def space_to_hx_index(self, x, y, z):
        return self.space_to_index(const.Hx, x, y, z)

    def space_to_hy_index(self, x, y, z):
        return self.space_to_index(const.Hy, x, y, z)

    def space_to_hz_index(self, x, y, z):
        return self.space_to_index(const.Hz, x, y, z)

    def display_info(self, indent=0):
        print(" " * indent, "Cartesian space")
        print(" " * indent, "MPI topology:", self.my_id, "of", self.numprocs)
        print(" " * indent, "size:", 2 * self.half_size, "resolution:", self.res)
        print(" " * indent, "dx:", self.dr[0], "dy:", self.dr[1], "dz:", self.dr[2])
        print(" " * indent, "number of participating nodes:", self.numprocs)


def in_range(idx, shape, component):
    """Bounds check of a storage index (gmes/geometry.py:751-812): the ghost
    planes (E: last transverse planes, H: plane 0) are out of range."""
    lo = [0, 0, 0]
    hi = [shape[0], shape[1], shape[2]]
    own = component.index % 3
    for a in range(3):
        if a == own:
            continue
        if component in const.E_COMPONENTS:
            hi[a] = shape[a] - 1
        elif component in const.H_COMPONENTS:
            lo[a] = 1
        else:
            raise ValueError
    if component not in const.COMPONENTS:
        raise ValueError
    for a in range(3):
        if idx[a] < lo[a] or idx[a] >= hi[a]:
            return False
    return True
