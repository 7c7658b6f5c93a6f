"""Tile builder: geometry list -> per-component material-ID grids.

Replaces the reference's per-point mapping loop ``init_material_<c>``
(gmes/fdtd.py:353-503), which asks ``GeomBoxTree.material_of_point`` for
every element of every component and attaches a C++ updater per cell.
Here every geometric object is tested once for a block of points (its
bounding box widened by one cell), giving the same (object, underneath)
decision per staggered point:

* owner = last object in the list that contains the point, else object 0
  (src/pygeom.pyx:143-156);
* for a compound (PML) owner, underneath = the owner before it was
  assigned, i.e. the last earlier object containing the point
  (src/pygeom.pyx:425-441);
* the ghost planes become Dummy whatever the geometry
  (gmes/fdtd.py:365-366,386-387,407-408,428-429,449-450,470-471).

Distinct (material, underneath) pairs become records of the engine's
table; CPML/UPML media contribute per-axis coefficient profiles evaluated
by the reference's scalar formulas at every storage coordinate.
"""

import numpy as np

from . import constant as const
from ._abi import CLS_DUMMY, MAT_REC_DTYPE
from .material import Dummy, Pml
from .pygeom import Compound, Shell


class TileSet(object):
    """Records, coefficient pool, per-component id grids and profiles."""

    def __init__(self):
        self.records = []          # list of dicts
        self._rec_index = {}
        self.coef = []
        self.ids = {}              # comp -> uint16 array (component shape)
        self.profiles = {}         # (prof, comp, axis) -> float64 [n][w]
        self.pml_sets = {}         # id(pml) -> prof index
        self.class_counts = {}     # comp -> {cls: count}
        self.type_counts = {}      # comp -> {reference updater prefix: count}
        self.owner_materials = {}  # comp -> {record id: material}

    def record_id(self, tile, pml_index):
        coef = tuple(float(v) for v in tile.get("coef", ()))
        value = tuple(float(v) for v in tile.get("value", (0.0, 0.0)))
        key = (tile["cls"], float(tile["eps_mu"]), value, pml_index, coef,
               tile.get("n1", 0), tile.get("n2", 0))
        rid = self._rec_index.get(key)
        if rid is None:
            rid = len(self.records)
            if rid > 65535:
                raise ValueError("more than 65536 distinct material records")
            self._rec_index[key] = rid
            off = len(self.coef)
            self.coef.extend(coef)
            self.records.append(dict(cls=tile["cls"], n1=tile.get("n1", 0), n2=tile.get("n2", 0),
                                     prof=pml_index if pml_index is not None else 0,
                                     coef_off=off, eps_mu=float(tile["eps_mu"]), value=value))
        return rid

    def record_array(self):
        arr = np.zeros(len(self.records), MAT_REC_DTYPE)
        for i, r in enumerate(self.records):
            arr[i] = (r["cls"], r["n1"], r["n2"], r["prof"], r["coef_off"], r["eps_mu"], r["value"])
        return arr

    def coef_array(self):
        return np.array(self.coef if self.coef else [0.0], np.float64)


def _window(coords, low, high, margin):
    """Index slice of sorted ``coords`` within [low - margin, high + margin]."""
    if not np.isfinite(low) and not np.isfinite(high):
        return slice(0, len(coords))
    lo = 0 if not np.isfinite(low) else int(np.searchsorted(coords, low - margin, "left"))
    hi = len(coords) if not np.isfinite(high) else int(np.searchsorted(coords, high + margin, "right"))
    return slice(lo, max(lo, hi))


def _object_mask(obj, coords, dr):
    """(slices, mask) of the points of a block where obj.in_object holds."""
    if isinstance(obj, Shell):
        full = tuple(slice(0, len(c)) for c in coords)
        mask = np.zeros(tuple(len(c) for c in coords), bool)
        for box in obj.box_list:
            w = tuple(_window(coords[a], box.low[a], box.high[a], dr[a]) for a in range(3))
            X = coords[0][w[0]][:, None, None]
            Y = coords[1][w[1]][None, :, None]
            Z = coords[2][w[2]][None, None, :]
            mask[w] |= box._mask(X, Y, Z)
        return full, mask
    box = obj.box
    w = tuple(_window(coords[a], box.low[a], box.high[a], dr[a]) for a in range(3))
    X = coords[0][w[0]][:, None, None]
    Y = coords[1][w[1]][None, :, None]
    Z = coords[2][w[2]][None, None, :]
    return w, np.asarray(obj._mask(X, Y, Z), bool)


def owner_map(space, comp, shape, geom_list):
    """Owner and underneath object index of every point of ``comp``."""
    coords = [space.axis_coords(comp, a, shape[a]) for a in range(3)]
    owner = np.zeros(shape, np.int32)
    under = np.zeros(shape, np.int32)
    for o in range(1, len(geom_list)):
        obj = geom_list[o]
        w, m = _object_mask(obj, coords, space.dr)
        if not m.any():
            continue
        if isinstance(obj.material, Compound):
            sub_u = under[w]
            sub_u[m] = owner[w][m]
            under[w] = sub_u
        sub_o = owner[w]
        sub_o[m] = o
        owner[w] = sub_o
    return owner, under, coords


def ghost_mask_slices(comp, shape):
    """Slices of the Dummy ghost planes of ``comp``."""
    own = comp.index % 3
    out = []
    for a in range(3):
        if a == own:
            continue
        sl = [slice(None)] * 3
        sl[a] = shape[a] - 1 if comp in const.E_COMPONENTS else 0
        out.append(tuple(sl))
    return out


# The reference names an updater type after the material class that makes
# it (src/material.pyx get_pw_material_<c>: DrudeHxReal, CpmlEzCmplx, ...);
# subclasses reuse their base's updaters (DcpRc makes DcpPlrc updaters).
_UPDATER_PREFIXES = ("Dummy", "Const", "Dielectric", "Upml", "Cpml", "DcpAde", "DcpPlrc", "Drude",
                     "Lorentz", "Dm2")


def updater_prefix(material):
    for klass in type(material).__mro__:
        if klass.__name__ in _UPDATER_PREFIXES:
            return klass.__name__
    return type(material).__name__


def build_tiles(space, geom_list, e_comps, h_comps, cmplx):
    """Map every active component; returns a TileSet."""
    ts = TileSet()
    dummy_rid = ts.record_id(Dummy()._tile(True, None, cmplx), None)
    for comp in tuple(e_comps) + tuple(h_comps):
        electric = comp in const.E_COMPONENTS
        shape = space.storage_shape(comp, tuple(e_comps) + tuple(h_comps))
        owner, under, coords = owner_map(space, comp, shape, geom_list)
        compound = np.array([isinstance(g.material, Compound) for g in geom_list])
        nkey = len(geom_list) + 1
        key = owner * nkey
        comp_owner = compound[owner]
        key[comp_owner] += under[comp_owner] + 1
        present = np.flatnonzero(np.bincount(key.ravel(), minlength=nkey * nkey))
        lut = np.zeros(nkey * nkey, np.uint16)
        mats = {}
        for kv in present:
            o, u = divmod(int(kv), nkey)
            mat = geom_list[o].material
            underneath = geom_list[u - 1].material if u > 0 else None
            tile = mat._tile(electric, underneath, cmplx)
            pml_index = None
            if isinstance(tile.get("pml"), Pml):
                pml = tile["pml"]
                pml_index = ts.pml_sets.setdefault(id(pml), len(ts.pml_sets))
                for a in range(3):
                    kk = (pml_index, comp.index, a)
                    if kk not in ts.profiles:
                        ts.profiles[kk] = pml.profile(coords[a], a)
            rid = ts.record_id(tile, pml_index)
            lut[kv] = rid
            mats[rid] = mat
        ids = lut[key]
        ghost = np.zeros(shape, bool)
        for sl in ghost_mask_slices(comp, shape):
            ids[sl] = dummy_rid
            ghost[sl] = True
        net = np.bincount(key[~ghost], minlength=nkey * nkey)
        types = {"Dummy": int(ghost.sum())} if ghost.any() else {}
        for kv in present:
            if net[kv]:
                name = updater_prefix(geom_list[int(kv) // nkey].material)
                types[name] = types.get(name, 0) + int(net[kv])
        ts.type_counts[comp] = types
        ts.ids[comp] = np.ascontiguousarray(ids)
        cls_of = np.array([r["cls"] for r in ts.records], np.int32)
        counts = np.bincount(cls_of[ids.ravel()], minlength=9)
        ts.class_counts[comp] = {c: int(v) for c, v in enumerate(counts) if v}
        ts.owner_materials[comp] = mats
    if len(ts.pml_sets) > 64:
        raise ValueError("at most 64 distinct PML media per simulation")
    return ts
