"""Batched assembly on the GPU: many uneven 2D problems (sizes, orders,
distortion, coefficient fields) assembled by one persistent launch captured
in a HIP graph, each problem's matrix checked against the oracle (pattern
bit-exact, FP64 values within relative Frobenius error 1e-13, mixed within
1e-6) and against its own single-problem assembly; the same for batches of
uneven 3D problems (Q1/Q2 hexahedra) and of P1/P2 triangles, each cell kind
an instantiation of the batched kernel of its own."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-13
TOL_MIXED = 1e-6


def relfro(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def make_problems(GT, O, count, seed, nmin=2, nmax=14, with_oracle=True):
    u = GT.inputs.uniform(seed, 4 * count)
    out = []
    for i in range(count):
        nx = nmin + int(u[4 * i] * (nmax - nmin + 1))
        ny = nmin + int(u[4 * i + 1] * (nmax - nmin + 1))
        order = 1 + int(u[4 * i + 2] * 3)
        perturb = 0.2 if u[4 * i + 3] < 0.5 else 0.0
        m = GT.cartesian_mesh((0.0, 1.0, 0.0, 1.0), (nx, ny))
        if perturb:
            m.coords = GT.perturb_interior_nodes(m, perturb, seed + i)
        Om = GT.interior(m)
        V = GT.lagrange_space(Om, order, dirichlet_boundary=GT.boundary(m))
        dO = GT.measure(Om, 2 * order)
        kappa = GT.inputs.lognormal(seed + 1000 + i, m.cells.shape[0])
        p = None
        if with_oracle:
            p = O.Problem((nx, ny)).space(order, dirichlet=1)
            p.set_coords(m.coords)
            p.quadrature(2 * order)
        out.append(dict(m=m, Om=Om, V=V, dO=dO, kappa=kappa, p=p, order=order))
    return out


def loads_of(GT, probs, seed):
    """A load per problem, cycling through the three source kinds: seeded
    Gaussian nodal values (free then Dirichlet), per-cell values, a
    constant.  Returns the forms and the oracle's (kind, data)."""
    forms, ref = [], []
    for i, q in enumerate(probs):
        V, dO = q["V"], q["dO"]
        k = i % 3
        if k == 0:
            f = GT.inputs.normal(seed + i, V.nfree + V.ndiri)
            forms.append(GT.Load(dO, GT.DiscreteField(V, f[:V.nfree], f[V.nfree:])))
            ref.append((2, f))
        elif k == 1:
            f = GT.inputs.uniform(seed + i, q["m"].cells.shape[0]) - 0.5
            forms.append(GT.Load(dO, GT.face_constant_field(f, q["Om"])))
            ref.append((1, f))
        else:
            forms.append(GT.Load(dO, 1.5))
            ref.append((0, np.array([1.5])))
    return forms, ref


def forms_of(GT, probs, kind, scale=1.0):
    F = GT.Laplace if kind == "laplace" else GT.Mass
    return [F(q["dO"], GT.face_constant_field(scale * q["kappa"], q["Om"])) for q in probs]


@pytest.mark.parametrize("kind", ["laplace", "mass"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_matches_oracle(GT, O, kind, precision):
    print(f"case batch {kind} {precision} 12 problems", flush=True)
    probs = make_problems(GT, O, 12, seed=41)
    prec = GT.FP64 if precision == "fp64" else GT.MIXED
    forms = forms_of(GT, probs, kind)
    loads, lref = loads_of(GT, probs, 900)
    b = GT.MatrixBatch(forms, [q["V"] for q in probs], precision=prec, tuning=None, loads=loads)
    b.run()
    mats = b.matrices()
    tol = TOL_FP64 if precision == "fp64" else TOL_MIXED
    code = O.LAPLACE if kind == "laplace" else O.MASS
    for q, A in zip(probs, mats):
        cp, rv, nz = q["p"].assemble_matrix(code, coef=q["kappa"])
        assert np.array_equal(A.colptr, cp) and np.array_equal(A.rowval, rv)
        assert relfro(A.nzval, nz) <= tol, (q["order"], q["m"].cells.shape)
    # load vectors stay FP64 in both precisions
    for q, v, (fk, fd) in zip(probs, b.vectors(), lref):
        assert relfro(v, q["p"].assemble_vector(fk, fd)) <= TOL_FP64, (q["order"], fk)
    # the same problems one call each: identical per-cell arithmetic and
    # cell-ascending sums per nonzero, so the values agree to rounding
    for f, q, A in zip(forms, probs, mats):
        A1 = GT.assemble_matrix(f, float, q["V"], q["V"], precision=prec)
        assert relfro(A.nzval, A1.nzval) <= 1e-15
    b.close()


def test_batch_tile_sizes_and_replays(GT, O):
    """Small tiles split every problem over many work items; the values do
    not depend on the tiling (each nonzero sums its cells in ascending
    order), and every replay of the graph gives the same values, with
    unrelated device work between the replays."""
    print("case batch tiles/replays", flush=True)
    probs = make_problems(GT, O, 10, seed=5, nmin=4, nmax=18, with_oracle=False)
    spaces = [q["V"] for q in probs]
    forms = forms_of(GT, probs, "laplace")
    b0 = GT.MatrixBatch(forms, spaces, precision=GT.MIXED, tuning=None)
    b0.run()
    ref = [A.nzval.copy() for A in b0.matrices()]
    b0.close()
    b1 = GT.MatrixBatch(forms, spaces, precision=GT.MIXED, tuning={1: 24, 2: 24, 3: 24, "waves": 12})
    other = GT.assemble_matrix(GT.Mass(probs[0]["dO"]), float, spaces[0], spaces[0])  # eager work on another form
    for rep in range(4):
        b1.run()
        vals = [A.nzval for A in b1.matrices()]
        for r, v in zip(ref, vals):
            assert np.array_equal(r, v), rep
        GT.assemble_vector(GT.Load(probs[1]["dO"], 1.0), float, spaces[1])
    assert other.nnz > 0
    b1.close()


def test_batch_new_coefficients(GT, O):
    """set_forms: new per-cell coefficients, the same graph runs again."""
    print("case batch set_forms", flush=True)
    probs = make_problems(GT, O, 6, seed=77)
    spaces = [q["V"] for q in probs]
    loads, _ = loads_of(GT, probs, 300)
    b = GT.MatrixBatch(forms_of(GT, probs, "laplace"), spaces, precision=GT.FP64, tuning=None, loads=loads)
    b.run()
    b.set_forms(forms_of(GT, probs, "laplace", scale=2.5))
    loads, lref = loads_of(GT, probs, 301)  # same kinds, new values: no rebuild
    b.set_loads(loads)
    b.run()
    for q, A in zip(probs, b.matrices()):
        cp, rv, nz = q["p"].assemble_matrix(O.LAPLACE, coef=2.5 * q["kappa"])
        assert relfro(A.nzval, nz) <= TOL_FP64
    for q, v, (fk, fd) in zip(probs, b.vectors(), lref):
        assert relfro(v, q["p"].assemble_vector(fk, fd)) <= TOL_FP64
    # a new kind of load (constant everywhere) rebuilds the work items
    b.set_loads([GT.Load(q["dO"], -2.0) for q in probs])
    b.run()
    for q, v in zip(probs, b.vectors()):
        assert relfro(v, q["p"].assemble_vector(0, [-2.0])) <= TOL_FP64
    # kappa reset to 1 (no coefficient buffer): picked up as well
    b.set_forms([GT.Laplace(q["dO"]) for q in probs])
    b.run()
    for q, A in zip(probs, b.matrices()):
        cp, rv, nz = q["p"].assemble_matrix(O.LAPLACE)
        assert relfro(A.nzval, nz) <= TOL_FP64
    b.close()


def test_batch_autotune_cache(GT, O, tmp_path):
    print("case batch autotune", flush=True)
    probs = make_problems(GT, O, 9, seed=13, nmin=6, nmax=24, with_oracle=False)
    spaces = [q["V"] for q in probs]
    forms = forms_of(GT, probs, "laplace")
    cache = str(tmp_path / "tiles.json")
    b = GT.MatrixBatch(forms, spaces, precision=GT.MIXED, tuning="auto", cache=cache)
    rep = b.tuning_report
    orders = {q["order"] for q in probs}
    types = [GT.batch.ELEM_NAMES[GT.batch.ELEM_TYPES[(GT.refcell.QUAD, o)]] for o in orders]
    assert all(not rep[t]["cached"] for t in types)
    for t in types:
        assert rep[t]["max_cells"] in GT.batch.TILE_CANDIDATES
    assert rep["waves"] in GT.batch.WAVE_CANDIDATES and b.waves == rep["waves"]
    b.run()
    vals = [A.nzval.copy() for A in b.matrices()]
    b.close()
    b2 = GT.MatrixBatch(forms, spaces, precision=GT.MIXED, tuning="auto", cache=cache)
    assert all(b2.tuning_report[t]["cached"] for t in types)
    assert b2.tiles == b.tiles and b2.waves == b.waves
    b2.run()
    for v, A in zip(vals, b2.matrices()):
        assert np.array_equal(v, A.nzval)
    b2.close()


def test_batch_rejects(GT, O):
    m = GT.cartesian_mesh((0.0, 1.0) * 3, (2, 2, 2), simplexify=True)  # tetrahedra: no batched kernel
    V = GT.lagrange_space(GT.interior(m), 1)
    with pytest.raises(NotImplementedError):
        GT.MatrixBatch([GT.Laplace(GT.measure(GT.interior(m), 2))], [V], tuning=None)
    m2 = GT.cartesian_mesh((0.0, 1.0) * 2, (3, 3))
    V2 = GT.lagrange_space(GT.interior(m2), 2)
    dO = GT.measure(GT.interior(m2), 4)
    with pytest.raises(NotImplementedError):
        GT.MatrixBatch([GT.Laplace(dO), GT.Mass(dO)], [V2, V2], tuning=None)
    with pytest.raises(NotImplementedError):  # not (order+1)^2 points
        GT.MatrixBatch([GT.Laplace(GT.measure(GT.interior(m2), 6))], [V2], tuning=None)


def make_problems_3d(GT, O, count, seed, nmin=2, nmax=6):
    u = GT.inputs.uniform(seed, 5 * count)
    out = []
    for i in range(count):
        cells = tuple(nmin + int(u[5 * i + d] * (nmax - nmin + 1)) for d in range(3))
        order = 1 + int(u[5 * i + 3] * 2)
        m = GT.cartesian_mesh((0.0, 1.0) * 3, cells)
        if u[5 * i + 4] < 0.5:
            m.coords = GT.perturb_interior_nodes(m, 0.15, seed + i)
        Om = GT.interior(m)
        V = GT.lagrange_space(Om, order, dirichlet_boundary=GT.boundary(m))
        dO = GT.measure(Om, 2 * order)
        kappa = GT.inputs.lognormal(seed + 1000 + i, m.cells.shape[0])
        p = O.Problem(list(cells)).space(order, dirichlet=1)
        p.set_coords(m.coords)
        p.quadrature(2 * order)
        out.append(dict(m=m, Om=Om, V=V, dO=dO, kappa=kappa, p=p, order=order))
    return out


@pytest.mark.parametrize("kind", ["laplace", "mass"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_3d_matches_oracle(GT, O, kind, precision):
    print(f"case batch 3d {kind} {precision} 8 problems", flush=True)
    probs = make_problems_3d(GT, O, 8, seed=61)
    assert {q["order"] for q in probs} == {1, 2}
    prec = GT.FP64 if precision == "fp64" else GT.MIXED
    forms = forms_of(GT, probs, kind)
    loads, lref = loads_of(GT, probs, 700)
    tol = TOL_FP64 if precision == "fp64" else TOL_MIXED
    code = O.LAPLACE if kind == "laplace" else O.MASS
    for tuning in (None, {1: 16, 2: 16, "waves": 8}):
        b = GT.MatrixBatch(forms, [q["V"] for q in probs], precision=prec, tuning=tuning, loads=loads)
        b.run()
        for q, A in zip(probs, b.matrices()):
            cp, rv, nz = q["p"].assemble_matrix(code, coef=q["kappa"])
            assert np.array_equal(A.colptr, cp) and np.array_equal(A.rowval, rv)
            assert relfro(A.nzval, nz) <= tol, (q["order"], q["m"].cells.shape, tuning)
        for q, v, (fk, fd) in zip(probs, b.vectors(), lref):
            assert relfro(v, q["p"].assemble_vector(fk, fd)) <= TOL_FP64, (q["order"], fk)
        b.close()


def test_batch_3d_waves_and_kinds(GT, O):
    """A 3D batch at 12 waves gives the 8-wave values bit for bit, and a
    batch mixing quadrilaterals and hexahedra is refused."""
    print("case batch 3d waves / mixed kinds", flush=True)
    probs = make_problems_3d(GT, O, 5, seed=19)
    spaces = [q["V"] for q in probs]
    forms = forms_of(GT, probs, "laplace")
    vals = []
    for w in (8, 12):
        b = GT.MatrixBatch(forms, spaces, precision=GT.MIXED, tuning={"waves": w})
        b.run()
        vals.append([A.nzval.copy() for A in b.matrices()])
        b.close()
    for a, c in zip(*vals):
        assert np.array_equal(a, c)
    q2 = make_problems(GT, O, 1, seed=3, with_oracle=False)[0]
    with pytest.raises(NotImplementedError):
        GT.MatrixBatch(forms[:1] + forms_of(GT, [q2], "laplace"), spaces[:1] + [q2["V"]], tuning=None)


def test_batch_3d_one_order_and_replays(GT, O):
    """A hexahedral batch runs one launch per element type: batches of Q1
    problems only and of Q2 problems only (one launch empty) match the
    oracle, and graph replays of a mixed batch (both launches, each with its
    own claim counter, cleared by the graph's first node) repeat the first
    run's values bit for bit."""
    print("case batch 3d one order / replays", flush=True)
    probs = make_problems_3d(GT, O, 8, seed=61)
    loads, lref = loads_of(GT, probs, 900)
    for order in (1, 2):
        sel = [i for i, q in enumerate(probs) if q["order"] == order]
        assert len(sel) >= 2
        sub = [probs[i] for i in sel]
        b = GT.MatrixBatch(forms_of(GT, sub, "laplace"), [q["V"] for q in sub], precision=GT.MIXED, tuning=None,
                           loads=[loads[i] for i in sel])
        b.run()
        for q, A in zip(sub, b.matrices()):
            cp, rv, nz = q["p"].assemble_matrix(O.LAPLACE, coef=q["kappa"])
            assert np.array_equal(A.colptr, cp) and np.array_equal(A.rowval, rv)
            assert relfro(A.nzval, nz) <= TOL_MIXED, (order, q["m"].cells.shape)
        for i, v in zip(sel, b.vectors()):
            fk, fd = lref[i]
            assert relfro(v, probs[i]["p"].assemble_vector(fk, fd)) <= TOL_FP64, (order, fk)
        b.close()
    b = GT.MatrixBatch(forms_of(GT, probs, "laplace"), [q["V"] for q in probs], precision=GT.MIXED, tuning=None,
                       loads=loads)
    b.run()
    first = [A.nzval.copy() for A in b.matrices()]
    vfirst = [v.copy() for v in b.vectors()]
    for _ in range(3):
        b.run()
        for a, A in zip(first, b.matrices()):
            assert np.array_equal(a, A.nzval)
        for a, v in zip(vfirst, b.vectors()):
            assert np.array_equal(a, v)
    b.close()


def make_problems_tri(GT, O, count, seed, nmin=2, nmax=12):
    u = GT.inputs.uniform(seed, 4 * count)
    out = []
    for i in range(count):
        cells = (nmin + int(u[4 * i] * (nmax - nmin + 1)), nmin + int(u[4 * i + 1] * (nmax - nmin + 1)))
        order = 1 + int(u[4 * i + 2] * 2)
        m = GT.cartesian_mesh((0.0, 1.0) * 2, cells, simplexify=True)
        if u[4 * i + 3] < 0.5:
            m.coords = GT.perturb_interior_nodes(m, 0.15, seed + i)
        Om = GT.interior(m)
        V = GT.lagrange_space(Om, order, dirichlet_boundary=GT.boundary(m))
        dO = GT.measure(Om, 2 * order)
        kappa = GT.inputs.lognormal(seed + 1000 + i, m.cells.shape[0])
        p = O.Problem(list(cells), simplexify=True).space(order, dirichlet=1)
        p.set_coords(m.coords)
        p.quadrature(2 * order)
        out.append(dict(m=m, Om=Om, V=V, dO=dO, kappa=kappa, p=p, order=order))
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_triangles_match_oracle(GT, O, precision):
    """A batch of uneven P1/P2 triangle problems (the batched kernel's
    triangle instantiation), Laplace with lognormal kappa and loads."""
    print(f"case batch triangles {precision}", flush=True)
    probs = make_problems_tri(GT, O, 8, seed=29)
    assert {q["order"] for q in probs} == {1, 2}
    prec = GT.FP64 if precision == "fp64" else GT.MIXED
    tol = TOL_FP64 if precision == "fp64" else TOL_MIXED
    forms = forms_of(GT, probs, "laplace")
    loads, lref = loads_of(GT, probs, 500)
    b = GT.MatrixBatch(forms, [q["V"] for q in probs], precision=prec, tuning=None, loads=loads)
    b.run()
    for q, A in zip(probs, b.matrices()):
        cp, rv, nz = q["p"].assemble_matrix(O.LAPLACE, coef=q["kappa"])
        assert np.array_equal(A.colptr, cp) and np.array_equal(A.rowval, rv)
        assert relfro(A.nzval, nz) <= tol, (q["order"], q["m"].cells.shape)
    for q, v, (fk, fd) in zip(probs, b.vectors(), lref):
        assert relfro(v, q["p"].assemble_vector(fk, fd)) <= TOL_FP64, (q["order"], fk)
    b.close()
