"""Geometric primitives and the point -> material lookup.

Drop-in counterpart of the reference's Cython module ``gmes.pygeom``
(src/pygeom.pyx).  Each primitive keeps the reference's inclusion test
(``in_object``) operation for operation, and adds ``_mask`` -- the same
test evaluated for a whole block of grid points -- which the tile builder
uses instead of the reference's per-point Python loop
(gmes/fdtd.py:353-477).

``GeomBoxTree.object_of_point`` returns the last object of the geometry
list whose ``in_object`` holds (the first object when none does), and the
"underneath" object of a compound (PML) material.  The reference reaches
the same answer through a bounding-box search tree
(src/pygeom.pyx:143-156,300-454): any object containing a point has a
bounding box that overlaps the point's leaf, and leaf lists keep list
order, so the tree only prunes objects that cannot contain the point.
"""

from copy import deepcopy
from math import sqrt

import numpy as np


def _norm(p):
    """src/pygeom.pyx:15-16"""
    return sqrt(p[0] ** 2 + p[1] ** 2 + p[2] ** 2)


class Material(object):
    """Base class of material types (src/pygeom.pyx:19-117)."""

    def __init__(self, eps_inf=1, mu_inf=1):
        self.eps_inf = float(eps_inf)
        self.mu_inf = float(mu_inf)

    def __getstate__(self):
        return dict(self.__dict__)

    def __setstate__(self, d):
        self.__dict__.update(d)

    def display_info(self, indent=0):
        raise NotImplementedError

    def init(self, space, param=None):
        raise NotImplementedError


class Compound(object):
    """Marker for materials that refer to an underlying medium
    (src/pygeom.pyx:120-124)."""


class GeomBox(object):
    """Axis-aligned bounding box, inclusive on both ends
    (src/pygeom.pyx:159-255)."""

    def __init__(self, low=(0, 0, 0), high=(0, 0, 0)):
        self.low = np.array(low, np.double)
        self.high = np.array(high, np.double)

    def union(self, box):
        self.low[:] = [min(a, b) for a, b in zip(self.low, box.low)]
        self.high[:] = [max(a, b) for a, b in zip(self.high, box.high)]

    def intersection(self, box):
        self.low[:] = [max(a, b) for a, b in zip(self.low, box.low)]
        self.high[:] = [min(a, b) for a, b in zip(self.high, box.high)]

    def add_point(self, point):
        self.low[:] = [min(a, b) for a, b in zip(self.low, point)]
        self.high[:] = [max(a, b) for a, b in zip(self.high, point)]

    @staticmethod
    def between(x, low, high):
        return low <= x <= high

    def in_box(self, point):
        return bool(self.between(point[0], self.low[0], self.high[0])
                    and self.between(point[1], self.low[1], self.high[1])
                    and self.between(point[2], self.low[2], self.high[2]))

    def _mask(self, X, Y, Z):
        return ((self.low[0] <= X) & (X <= self.high[0]) & (self.low[1] <= Y)
                & (Y <= self.high[1]) & (self.low[2] <= Z) & (Z <= self.high[2]))

    def overlap(self, box):
        b = self.between
        return ((b(box.low[0], self.low[0], self.high[0]) or b(box.high[0], self.low[0], self.high[0])
                 or b(self.low[0], box.low[0], box.high[0]))
                and (b(box.low[1], self.low[1], self.high[1]) or b(box.high[1], self.low[1], self.high[1])
                     or b(self.low[1], box.low[1], box.high[1]))
                and (b(box.low[2], self.low[2], self.high[2]) or b(box.high[2], self.low[2], self.high[2])
                     or b(self.low[2], box.low[2], box.high[2])))

    def divide(self, axis, x):
        high1 = list(self.high)
        high1[axis] = x
        low2 = list(self.low)
        low2[axis] = x
        return GeomBox(self.low, high1), GeomBox(low2, self.high)

    def display_info(self, indent=0):
        print(" " * indent, "geom box:", "low:", self.low, "high:", self.high)

    def __str__(self):
        return "low: " + self.low.__str__() + " high: " + self.high.__str__()


class GeometricObject(object):
    """Base class of geometric objects (src/pygeom.pyx:470-535)."""

    def __init__(self, material):
        self.material = material
        self.box = None

    def init(self, space):
        self.material.init(space)
        self.box = self.geom_box()

    def geom_box(self):
        raise NotImplementedError

    def in_object(self, point):
        raise NotImplementedError

    def _mask(self, X, Y, Z):
        """``in_object`` for broadcastable coordinate arrays X, Y, Z."""
        raise NotImplementedError

    def display_info(self, indent=0):
        print(" " * indent, "geometric object")
        if self.material:
            self.material.display_info(indent + 5)


class DefaultMedium(GeometricObject):
    """The whole space (src/pygeom.pyx:538-568)."""

    def in_object(self, point):
        return self.box.in_box(point)

    def _mask(self, X, Y, Z):
        return self.box._mask(X, Y, Z)

    def geom_box(self):
        return GeomBox((-np.inf, -np.inf, -np.inf), (np.inf, np.inf, np.inf))

    def display_info(self, indent=0):
        if self.material:
            self.material.display_info(indent + 5)


class Cone(GeometricObject):
    """A (possibly truncated) cone (src/pygeom.pyx:571-695)."""

    def __init__(self, material, center=(0, 0, 0), radius2=0, axis=(1, 0, 0), radius=1, height=1):
        GeometricObject.__init__(self, material)
        if radius < 0:
            raise ValueError("radius must be non-negative.")
        self.radius = float(radius)
        if radius2 < 0:
            raise ValueError("radius2 must be non-negative.")
        self.radius2 = float(radius2)
        self.center = np.array(center, np.double)
        self.axis = np.array(axis, np.double) / _norm(axis)
        self.height = float(height)

    def in_object(self, point):
        p = np.array(point, np.double)
        r = p - self.center
        proj = float(np.dot(self.axis, r))
        if abs(proj) <= .5 * self.height:
            if self.radius2 == self.radius == np.inf:
                return True
            radius = self.radius
            radius += (proj / self.height + .5) * (self.radius2 - radius)
            return _norm(r - proj * self.axis) <= abs(radius)
        return False

    def _mask(self, X, Y, Z):
        r0, r1, r2 = X - self.center[0], Y - self.center[1], Z - self.center[2]
        a = self.axis
        proj = (a[0] * r0 + a[1] * r1) + a[2] * r2
        inside = np.abs(proj) <= .5 * self.height
        if self.radius2 == self.radius == np.inf:
            return inside
        radius = self.radius + (proj / self.height + .5) * (self.radius2 - self.radius)
        q0, q1, q2 = r0 - proj * a[0], r1 - proj * a[1], r2 - proj * a[2]
        nrm = np.sqrt((q0 * q0 + q1 * q1) + q2 * q2)
        return inside & (nrm <= np.abs(radius))

    def geom_box(self):
        h = .5 * self.height
        radial = np.sqrt(np.maximum(0, 1 - self.axis * self.axis))
        extent1 = np.zeros((3,), np.double)
        extent2 = np.zeros((3,), np.double)
        np.multiply(self.radius, radial, out=extent1, where=radial != 0)
        np.multiply(self.radius2, radial, out=extent2, where=radial != 0)
        box1 = GeomBox(low=self.center, high=self.center)
        box1.low -= h * self.axis + extent1
        box1.high -= h * self.axis - extent1
        box2 = GeomBox(low=self.center, high=self.center)
        box2.low += h * self.axis - extent2
        box2.high += h * self.axis + extent2
        box1.union(box2)
        return box1


class Cylinder(Cone):
    """A cylinder (src/pygeom.pyx:698-722)."""

    def __init__(self, material, center=(0, 0, 0), axis=(0, 0, 1), radius=1, height=1):
        Cone.__init__(self, material, center, radius, axis, radius, height)


class Block(GeometricObject):
    """A parallelepiped (src/pygeom.pyx:725-810)."""

    def __init__(self, material, center=(0, 0, 0), e1=(1, 0, 0), e2=(0, 1, 0), e3=(0, 0, 1),
                 size=(1, 1, 1)):
        GeometricObject.__init__(self, material)
        self.center = np.array(center, np.double)
        self.e1 = np.array(e1, np.double) / _norm(e1)
        self.e2 = np.array(e2, np.double) / _norm(e2)
        self.e3 = np.array(e3, np.double) / _norm(e3)
        self.size = np.array(size, np.double)
        basis = np.column_stack((self.e1, self.e2, self.e3))
        self.projection_matrix = np.linalg.inv(basis)

    def in_object(self, point):
        p = np.array(point, np.double)
        r = p - self.center
        proj = np.dot(self.projection_matrix, r)
        return bool((np.abs(proj) <= .5 * self.size).all())

    def _proj(self, X, Y, Z):
        r0, r1, r2 = X - self.center[0], Y - self.center[1], Z - self.center[2]
        m = self.projection_matrix
        return [(m[i, 0] * r0 + m[i, 1] * r1) + m[i, 2] * r2 for i in range(3)]

    def _mask(self, X, Y, Z):
        pr = self._proj(X, Y, Z)
        half = .5 * self.size
        return (np.abs(pr[0]) <= half[0]) & (np.abs(pr[1]) <= half[1]) & (np.abs(pr[2]) <= half[2])

    def geom_box(self):
        box = GeomBox(low=self.center, high=self.center)
        s1 = self.size[0] * self.e1
        s2 = self.size[1] * self.e2
        s3 = self.size[2] * self.e3
        corner = self.center - 0.5 * (s1 + s2 + s3)
        for pt in (corner, corner + s1, corner + s2, corner + s3, corner + s1 + s2,
                   corner + s1 + s3, corner + s3 + s2, corner + s1 + s2 + s3):
            box.add_point(pt)
        return box


class Ellipsoid(Block):
    """An ellipsoid (src/pygeom.pyx:813-856)."""

    def __init__(self, material, center=(0, 0, 0), e1=(1, 0, 0), e2=(0, 1, 0), e3=(0, 0, 1),
                 size=(1, 1, 1)):
        Block.__init__(self, material, center, e1, e2, e3, size)
        self.inverse_semi_axes = 2 / np.array(size, np.double)

    def in_object(self, point):
        p = np.array(point, np.double)
        r = p - self.center
        proj = np.dot(self.projection_matrix, r)
        q = proj * self.inverse_semi_axes
        return bool(sum(q * q) <= 1)

    def _mask(self, X, Y, Z):
        pr = self._proj(X, Y, Z)
        q = [pr[i] * self.inverse_semi_axes[i] for i in range(3)]
        return ((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) <= 1


class Sphere(GeometricObject):
    """A sphere (src/pygeom.pyx:859-910)."""

    def __init__(self, material, center=(0, 0, 0), radius=1):
        GeometricObject.__init__(self, material)
        if radius < 0:
            raise ValueError("radius must be non-negative.")
        self.radius = float(radius)
        self.center = np.array(center, np.double)

    def geom_box(self):
        box = GeomBox(low=self.center, high=self.center)
        box.low -= self.radius
        box.high += self.radius
        return box

    def in_object(self, point):
        p = np.array(point, np.double)
        return _norm(p - self.center) <= self.radius

    def _mask(self, X, Y, Z):
        r0, r1, r2 = X - self.center[0], Y - self.center[1], Z - self.center[2]
        return np.sqrt((r0 * r0 + r1 * r1) + r2 * r2) <= self.radius


class Shell(GeometricObject):
    """A boundary shell made of up to six slabs (src/pygeom.pyx:913-1148)."""

    def __init__(self, material, center=(0, 0, 0), size=None, thickness=1, plus_x=True,
                 minus_x=True, plus_y=True, minus_y=True, plus_z=True, minus_z=True):
        GeometricObject.__init__(self, material)
        self.center = np.array(center, np.double)
        self.d = float(thickness)
        self.box_list = []
        self.minus_x, self.plus_x = minus_x, plus_x
        self.minus_y, self.plus_y = minus_y, plus_y
        self.minus_z, self.plus_z = minus_z, plus_z
        if size is None:
            self.half_size = np.zeros((3,), np.double)
            self.boundary = True
        else:
            self.half_size = np.array([0.5 * v for v in size], np.double)
            self.boundary = False

    def init(self, space):
        if self.boundary:
            self.half_size[:] = space.half_size
        for i in range(3):
            if 2 * self.half_size[i] < space.dr[i]:
                self.half_size[i] = 0.5 * space.dr[i]
        if 2 * self.half_size[0] <= space.dr[0]:
            self.plus_x = self.minus_x = False
        if 2 * self.half_size[1] <= space.dr[1]:
            self.plus_y = self.minus_y = False
        if 2 * self.half_size[2] <= space.dr[2]:
            self.plus_z = self.minus_z = False
        c, hs, d = self.center, self.half_size, self.d
        self.box_list = []
        if self.plus_x:
            self.box_list.append(GeomBox((c[0] + hs[0] - d, c[1] - hs[1], c[2] - hs[2]), c + hs))
        if self.minus_x:
            self.box_list.append(GeomBox(c - hs, (c[0] - hs[0] + d, c[1] + hs[1], c[2] + hs[2])))
        if self.plus_y:
            self.box_list.append(GeomBox((c[0] - hs[0], c[1] + hs[1] - d, c[2] - hs[2]), c + hs))
        if self.minus_y:
            self.box_list.append(GeomBox(c - hs, (c[0] + hs[0], c[1] - hs[1] + d, c[2] + hs[2])))
        if self.plus_z:
            self.box_list.append(GeomBox((c[0] - hs[0], c[1] - hs[1], c[2] + hs[2] - d), c + hs))
        if self.minus_z:
            self.box_list.append(GeomBox(c - hs, (c[0] + hs[0], c[1] + hs[1], c[2] - hs[2] + d)))
        self.material.init(space, (self.center, self.half_size, self.d))
        self.box = self.geom_box()

    def in_object(self, point):
        for box in self.box_list:
            if box.in_box(point):
                return True
        return False

    def _mask(self, X, Y, Z):
        out = None
        for box in self.box_list:
            m = box._mask(X, Y, Z)
            out = m if out is None else (out | m)
        if out is None:
            return np.zeros(np.broadcast(X, Y, Z).shape, bool)
        return out

    def geom_box(self):
        return GeomBox(self.center - self.half_size, self.center + self.half_size)


class GeomBoxTree(object):
    """Point -> (object, underneath object) lookup over a geometry list
    (src/pygeom.pyx:300-467; see the module docstring for why a linear
    scan from the end returns the tree's answer)."""

    def __init__(self, geom_list):
        self.geom_list = tuple(geom_list)

    @staticmethod
    def _find_object(point, geom_list):
        """src/pygeom.pyx:143-156: the last object containing point, else
        the first one."""
        i = len(geom_list) - 1
        while i > 0 and geom_list[i].in_object(point) is False:
            i -= 1
        return geom_list[i], i

    def object_of_point(self, point):
        point = tuple(float(v) for v in point)
        geom_obj, idx = self._find_object(point, self.geom_list)
        if isinstance(geom_obj.material, Compound):
            underneath_obj, _ = self._find_object(point, self.geom_list[:idx])
        else:
            underneath_obj = None
        return geom_obj, underneath_obj

    def material_of_point(self, point):
        geom_obj, underneath_obj = self.object_of_point(point)
        under = underneath_obj.material if underneath_obj is not None else None
        return geom_obj.material, under

    def display_info(self, node=None, indent=0):
        for g in self.geom_list:
            print(" " * indent, "bounding box:", g.box)
            g.display_info(indent + 5)
