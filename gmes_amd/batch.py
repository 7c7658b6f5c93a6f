"""Many simulations advanced together on one GPU.

A ``Batch`` groups initialised FDTD objects of one precision and field type
(real or complex).  Each keeps its own fields, materials, sources, probes
and clock; a step of the batch runs the main, source and ghost kernels of
every member as one launch each (the engine maps blocks to members through
a device job table) and is replayed from one captured HIP graph.  Members
may differ in size, mode and materials.  The batched main launches of 3-D
and collapsed-z members run flat blocks of 256 consecutive cells, so they
need no tile; autotune=True still tunes each 3-D member's tile for its own
launches (when it is stepped alone).

This has no counterpart in the reference, which advances one FDTD at a
time (gmes/fdtd.py:1099-1122); every member computes exactly what it would
compute alone.
"""


class Batch(object):
    def __init__(self, sims, autotune=False, graph=True):
        if not sims:
            raise ValueError("empty batch")
        self.sims = list(sims)
        lib = self.sims[0]._lib
        for s in self.sims:
            if s._sim is None:
                raise RuntimeError("init() must be called before batching")
            if s._lib is not lib:
                raise ValueError("batch members must use the same engine")
        self._lib = lib
        self.tiles = [lib.autotune(s._sim) if autotune else None for s in self.sims]
        self._b = lib.batch_create([s._sim for s in self.sims])
        lib.batch_set_graph(self._b, 1 if graph else 0)

    def advance(self, nsteps):
        """Advance every member by nsteps whole steps."""
        if nsteps <= 0:
            return
        if self._b is None:
            raise RuntimeError("batch is closed")
        for s in self.sims:
            s._prepare(nsteps)
        self._lib.batch_step(self._b, nsteps)
        for s in self.sims:
            s._after_steps(nsteps)

    def sync(self):
        self._lib.batch_sync(self._b)

    def member_errors(self):
        """Wait for the batch and return {member index: message} for the
        members whose kernels reported a deferred error (the DM2 corrector's
        non-convergence or invalid error, src/pw_dm2.hh:398-402).  The flag
        is per member, so the other members' results stand."""
        out = {}
        for i, s in enumerate(self.sims):
            try:
                self._lib.sync(s._sim)
            except RuntimeError as e:
                out[i] = str(e)
        return out

    def profile(self, enable):
        """Time the batch's kernel families with HIP events (see
        read_profile); steps run without the captured graph meanwhile."""
        self._lib.batch_profile(self._b, enable)

    def read_profile(self):
        """(ms, launches) per family: E 3-D main, H 3-D main, E 2-D main,
        H 2-D main, dispersive, the rest."""
        return self._lib.batch_read_profile(self._b)

    def close(self):
        """Dissolve the batch (the members stay usable on their own)."""
        if self._b is not None:
            self._lib.batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
