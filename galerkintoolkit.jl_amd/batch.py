"""Batched assembly: many independent problems of uneven size (meshes,
orders, coefficient fields) assembled by one persistent kernel launch (a
batch of hexahedra: one per element type), captured once in a HIP graph,
with the tile size of each element type picked by an autotuner that caches
its choices.

The reference assembles one problem per GT.assemble_matrix call
(src/problems.jl:319-350); MatrixBatch stands for a loop of such calls over
(form, space) pairs and returns the same SparseMatrixCSC per problem.
"""
import json
import os
import tempfile

import numpy as np

from . import _lib as L
from . import refcell as R
from .assembly import (FP64, FREE, Laplace, Mass, SparseMatrixCSC, _Assembler, _assembler, _coefficients,
                       _form_code, _load_data, _ncoef_items)

# element types of the batched kernels: scalar quadrilaterals of order 1, 2,
# 3 with (order+1)^2 Gauss points, scalar hexahedra of order 1, 2 with
# (order+1)^3, scalar triangles of order 1, 2 with degree-2k rules
# ((order+1)^2 points as well; a batch holds one cell kind)
ELEM_TYPES = {(R.QUAD, 1): 0, (R.QUAD, 2): 1, (R.QUAD, 3): 2, (R.HEX, 1): 3, (R.HEX, 2): 4, (R.TRI, 1): 6,
              (R.TRI, 2): 7}
ELEM_NAMES = {0: "quad-q1", 1: "quad-q2", 2: "quad-q3", 3: "hex-q1", 4: "hex-q2", 6: "tri-p1", 7: "tri-p2"}

# tile sizes the autotuner tries (cells per tile; 0 = as many as fit in LDS)
TILE_CANDIDATES = (0, 256, 160, 96, 64, 40)
# waves per workgroup it tries
WAVE_CANDIDATES = (8, 12)


def elem_type(space, dOmega=None):
    """Element type of the batched kernels for a space, or raise."""
    key = (space.mesh.cellkind, space.order)
    if space.ncomp != 1 or key not in ELEM_TYPES:
        raise NotImplementedError("batched problems are scalar Q1/Q2/Q3 Lagrange spaces on quadrilaterals, "
                                  "Q1/Q2 on hexahedra or P1/P2 on triangles")
    if dOmega is not None and len(dOmega.w) != (space.order + 1) ** space.mesh.D:
        raise NotImplementedError("batched problems use (order+1)^D Gauss points: GT.measure(Omega, 2*order)")
    return ELEM_TYPES[key]


def default_cache_path():
    p = os.environ.get("GTK_TUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "galerkintoolkit_jl_amd", "batch_tiles.json")


def _cache_key(form_code, precision, et):
    ver = L.load().gtk_version().decode()
    return f"{ver}|form{form_code}|prec{precision}|{ELEM_NAMES[et]}"


def _read_cache(path):
    try:
        with open(path) as f:
            d = json.load(f)
        return d if isinstance(d, dict) else {}
    except (OSError, ValueError):
        return {}


def _write_cache(path, d):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    fd, tmp = tempfile.mkstemp(dir=os.path.dirname(os.path.abspath(path)), suffix=".tmp")
    with os.fdopen(fd, "w") as f:
        json.dump(d, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _time_batch(batch, repeats):
    batch.run()  # first launch of a fresh graph
    return min(batch.run().kernel_ms for _ in range(repeats))


def autotune(problems, form_code, precision, device=0, cache=None, candidates=TILE_CANDIDATES,
             waves=WAVE_CANDIDATES, sample=8, repeats=3, retune=False):
    """Launch shape of a batch of `problems` (a list of (assembler,
    elem_type, cells)): waves per workgroup, and cells per tile for each
    element type present.  An element type without a cached entry is tuned
    on a sample of its problems (evenly spread over their sizes): every
    (waves, cells) candidate is built into a batch of the sample and timed by
    its kernel time per cell, and the best tile per wave count is cached under
    (library build, form, precision, element type).  The batch then takes the
    wave count with the least predicted time over its own cell mix.
    Returns (waves, {type: max_cells}, report)."""
    path = default_cache_path() if cache is None else cache
    table = _read_cache(path)
    report = {}
    entries = {}
    cells_of = {}
    for et in sorted({p[1] for p in problems}):
        mine = sorted((p for p in problems if p[1] == et), key=lambda p: p[2])
        cells_of[et] = sum(p[2] for p in mine)
        key = _cache_key(form_code, precision, et)
        e = table.get(key)
        if not retune and isinstance(e, dict) and all(str(w) in e for w in waves):
            entries[et] = e
            report[ELEM_NAMES[et]] = {"cached": True}
            continue
        pick = np.unique(np.linspace(0, len(mine) - 1, min(sample, len(mine))).round().astype(int))
        subset = [mine[i] for i in pick]
        cells = sum(p[2] for p in subset)
        e = {}
        for w in waves:
            times = {}
            for cap in candidates:
                b = L.Batch(device, form_code, precision)
                for a, _, _ in subset:
                    b.add(a.ctx)
                b.set_waves(w)
                b.set_tile_cells(et, cap)
                b.build()
                times[cap] = 1e6 * _time_batch(b, repeats) / cells
                b.close()
            best = min(times, key=lambda c: times[c])
            e[str(w)] = {"max_cells": best, "ns_per_cell": times[best],
                         "candidates": {str(c): round(t, 4) for c, t in times.items()}}
        e["sample_cells"] = cells
        table[key] = entries[et] = e
        report[ELEM_NAMES[et]] = {"cached": False}
    _write_cache(path, table)
    predicted = {w: sum(cells_of[et] * entries[et][str(w)]["ns_per_cell"] for et in entries) for w in waves}
    w = min(predicted, key=lambda k: predicted[k])
    tiles = {et: int(entries[et][str(w)]["max_cells"]) for et in entries}
    for et in entries:
        r = report[ELEM_NAMES[et]]
        r["max_cells"] = tiles[et]
        r["ns_per_cell"] = {str(k): round(entries[et][str(k)]["ns_per_cell"], 4) for k in waves}
    report["waves"] = w
    report["predicted_ms"] = {str(k): round(v * 1e-6, 4) for k, v in predicted.items()}
    return w, tiles, report


# launch shapes the single-problem tuner tries: waves per workgroup
# (matrix-core kernels; the VALU kernel: its default, 0, or 4) and tile caps
SINGLE_WAVES = {True: (8, 12, 16), False: (0, 4)}
SINGLE_CAPS = (0, 64)
_KIND_NAMES = {R.QUAD: "quad", R.HEX: "hex", R.TRI: "tri", R.TET: "tet"}


def element_name(space, dOmega):
    """Element type of a single-problem work plan: cell kind, order,
    components and quadrature points."""
    return f"{_KIND_NAMES[space.mesh.cellkind]}-k{space.order}-c{space.ncomp}-q{len(dOmega.w)}"


def autotune_matrix(form, space, precision=FP64, rows=FREE, cols=FREE, device=0, cache=None, retune=False,
                    repeats=3, launches=5):
    """Launch shape of one problem's (form, precision) work plan on the
    (rows, cols) block: waves per workgroup and a cap on the cells of a tile.
    For an element type (cell kind, order, components, quadrature points)
    without a cached entry, every candidate is built on this problem and its
    kernel timed (HIP events, best of `repeats` runs of `launches` launches);
    the fastest is cached under (library build, form, precision, element
    type) in the batch tuner's file and applied to the problem.  Later
    problems of the same type take the cached shape.  Returns a report."""
    code = _form_code(form)
    prec = L.GTK_FP64 if precision == FP64 else L.GTK_MIXED
    a = _assembler(space, form.dOmega, device)
    a.pattern(rows, cols)
    coef, ncoef = _coefficients(form, _ncoef_items(form, space))
    a.ctx.set_cell_coefficients(coef, ncoef)
    ver = L.load().gtk_version().decode()
    key = f"{ver}|single|form{code}|prec{prec}|{element_name(space, form.dOmega)}"
    path = default_cache_path() if cache is None else cache
    table = _read_cache(path)
    e = table.get(key)
    report = {"element": element_name(space, form.dOmega)}
    if retune or not (isinstance(e, dict) and "waves" in e and "max_cells" in e):
        mfma = bool(a.ctx.plan_info(code, prec, rows, cols)["matrix_cores"])
        times = {}
        for w in SINGLE_WAVES[mfma]:
            for cap in SINGLE_CAPS:
                try:
                    a.ctx.set_launch(code, prec, rows, cols, w, cap)
                except L.GtkError:
                    continue  # no such instantiation, or the tile does not fit
                a.ctx.time_matrix(code, prec, rows, cols, 1)
                times[(w, cap)] = min(a.ctx.time_matrix(code, prec, rows, cols, launches) for _ in range(repeats))
        best = min(times, key=lambda k: times[k])
        e = {"waves": best[0], "max_cells": best[1], "ms": times[best],
             "candidates": {f"{w}/{c}": round(t, 5) for (w, c), t in times.items()}}
        table[key] = e
        _write_cache(path, table)
        report["cached"] = False
        report["candidates_ms"] = e["candidates"]
    else:
        report["cached"] = True
    a.ctx.set_launch(code, prec, rows, cols, int(e["waves"]), int(e["max_cells"]))
    a.values_owner.pop((rows, cols), None)  # the timing launches replaced the block's device values
    report.update(waves=int(e["waves"]), max_cells=int(e["max_cells"]),
                  plan=a.ctx.plan_info(code, prec, rows, cols))
    return report


class MatrixBatch:
    """The (FREE, FREE) matrices of many problems, one bilinear form kind
    (GT Laplace or mass, with per-problem coefficients) and one precision,
    assembled together.

    forms[i] pairs with spaces[i]; forms[i].dOmega is that problem's
    GT.measure.  loads: optional GT Load forms, one per problem (or None
    entries), whose free-row vectors the same launch assembles.  tuning:
    "auto" (autotuner with its cache), a dict {order: max_cells}, or None
    (tiles as large as LDS allows)."""

    def __init__(self, forms, spaces, precision=FP64, device=0, tuning="auto", cache=None, loads=None):
        if len(forms) != len(spaces) or not forms:
            raise ValueError("one form per space, at least one problem")
        kinds = {type(f) for f in forms}
        if len(kinds) != 1 or not kinds <= {Laplace, Mass}:
            raise NotImplementedError("a batch holds one form kind, Laplace or Mass")
        self.form_code = _form_code(forms[0])
        self.precision = L.GTK_FP64 if precision == FP64 else L.GTK_MIXED
        self.device = device
        self.forms = list(forms)
        self.spaces = list(spaces)
        if len({V.mesh.cellkind for V in spaces}) != 1:
            raise NotImplementedError("a batch holds one cell kind")
        self.cellkind = spaces[0].mesh.cellkind
        self.problems = []
        for form, V in zip(forms, spaces):
            et = elem_type(V, form.dOmega)
            # a context of its own: coefficients set through single-problem
            # calls on the same space do not reach the batch
            a = _Assembler(V, form.dOmega, device)
            a.pattern(FREE, FREE)
            coef, ncoef = _coefficients(form, V.mesh.cells.shape[0])
            a.ctx.set_cell_coefficients(coef, ncoef)
            self.problems.append((a, et, V.mesh.cells.shape[0]))
        self._tune(tuning, cache)
        self.batch = L.Batch(device, self.form_code, self.precision)
        for a, _, _ in self.problems:
            self.batch.add(a.ctx)
        self.loads = [None] * len(self.problems)
        if loads is not None:
            if len(loads) != len(self.problems):
                raise ValueError("one load (or None) per problem")
            self.set_loads(loads)
        self.batch.set_waves(self.waves)
        for et, cap in self.tiles.items():
            self.batch.set_tile_cells(et, cap)
        self.batch.build()
        self.cells = sum(p[2] for p in self.problems)
        self.cache = cache
        self.tuning = tuning

    def _tune(self, tuning, cache):
        """tuning: "auto", None (8 waves, tiles as large as LDS allows) or a
        dict {order: max_cells, "waves": w} (orders of the batch's cell
        kind)."""
        self.tuning_report = None
        if tuning == "auto":
            self.waves, self.tiles, self.tuning_report = autotune(self.problems, self.form_code, self.precision,
                                                                  self.device, cache)
        elif tuning is None:
            self.waves, self.tiles = 8, {}
        else:
            t = dict(tuning)
            self.waves = int(t.pop("waves", 8))
            self.tiles = {ELEM_TYPES[(self.cellkind, o)]: int(c) for o, c in t.items()
                          if (self.cellkind, o) in ELEM_TYPES}

    def with_precision(self, precision, tuning=None):
        """A batch over the same problems and device contexts (patterns,
        coefficients, loads) in another precision.  Both write the same value
        arrays: read one's matrices before running the other."""
        other = object.__new__(MatrixBatch)
        other.__dict__.update(self.__dict__)
        other.precision = L.GTK_FP64 if precision == FP64 else L.GTK_MIXED
        other.forms, other.loads = list(self.forms), [None] * len(self.problems)
        other._tune(self.tuning if tuning is None else tuning, self.cache)
        other.batch = L.Batch(self.device, self.form_code, other.precision)
        for a, _, _ in self.problems:
            other.batch.add(a.ctx)
        if any(l is not None for l in self.loads):
            other.set_loads(self.loads)
        other.batch.set_waves(other.waves)
        for et, cap in other.tiles.items():
            other.batch.set_tile_cells(et, cap)
        other.batch.build()
        return other

    def set_forms(self, forms):
        """New coefficients for the same problems (same form kind); the next
        run assembles with them (the reference's update_matrix! per problem,
        src/assembly.jl:592-893)."""
        if len(forms) != len(self.problems) or {type(f) for f in forms} != {type(self.forms[0])}:
            raise ValueError("same number of problems and the same form kind")
        for form, (a, _, n) in zip(forms, self.problems):
            coef, ncoef = _coefficients(form, n)
            a.ctx.set_cell_coefficients(coef, ncoef)
        self.forms = list(forms)

    def set_loads(self, loads):
        """Load forms per problem (None: no vector); values of an existing
        kind take effect at the next run, a new kind rebuilds the batch."""
        rebuild = False
        for i, (l, V) in enumerate(zip(loads, self.spaces)):
            if l is None:
                if self.loads[i] is not None:
                    raise ValueError("a problem's load cannot be removed")
                continue
            if l.dOmega is not self.forms[i].dOmega:
                raise ValueError("a load integrates over its problem's measure")
            kind, data = _load_data(l, V)
            old = self.loads[i]
            rebuild |= old is None or _load_data(old, V)[0] != kind
            self.batch.set_load(i, kind, data)
            self.loads[i] = l
        if rebuild and getattr(self.batch, "built", False):
            self.batch.build()

    def run(self):
        """One launch assembling every problem; returns its gtk_stats."""
        return self.batch.run()

    def vectors(self):
        """The assembled free-row load vector of every problem with a load
        (None for the others), from the last run."""
        return [None if l is None else self.batch.get_vector(i, V.nfree)
                for i, (l, V) in enumerate(zip(self.loads, self.spaces))]

    def matrices(self):
        """The assembled SparseMatrixCSC of every problem (values copied
        from the device)."""
        out = []
        for a, _, _ in self.problems:
            nnz, ncols, nrows, colptr, rowval = a.pattern(FREE, FREE)
            A = SparseMatrixCSC(nrows, ncols, colptr, rowval, a.ctx.get_values(FREE, FREE, nnz))
            A.device = (a, FREE, FREE)
            out.append(A)
        return out

    def close(self):
        self.batch.close()


def assemble_matrices(forms, spaces, precision=FP64, device=0, tuning="auto", cache=None):
    """GT.assemble_matrix(forms[i], Float64, spaces[i], spaces[i]) for every
    i, in one launch; returns (matrices, batch) so the batch can be run again
    after new coefficients (MatrixBatch.set_forms)."""
    b = MatrixBatch(forms, spaces, precision=precision, device=device, tuning=tuning, cache=cache)
    b.run()
    return b.matrices(), b
