"""gtk_assemble_vector_noalloc through the C ABI: the load vector with no
device allocation (load values and vector staged in the library's
per-device scratch array) against gtk_assemble_vector, bit for bit, on two
calls in a row (the second on a smaller problem), then on the first problem
again and on a boundary load whose untouched rows the call must clear after
earlier calls left values in the scratch; plus the size limit and the
call-order check.  The vectors of gtk_assemble_vector are themselves checked
against the oracle here (relative error 1e-13)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def make(GT, cells, order, degree, diri, perturb, seed, boundary_measure=False):
    D = len(cells)
    m = GT.cartesian_mesh(tuple([0.0, 1.0] * D), cells)
    if perturb:
        m.coords = GT.perturb_interior_nodes(m, perturb, seed)
    Om = GT.interior(m)
    V = GT.lagrange_space(Om, order, dirichlet_boundary=GT.boundary(m) if diri else None)
    if boundary_measure:
        return m, V, GT.assembly._BoundaryAssembler(V, GT.measure(GT.boundary(m), degree))
    return m, V, GT.assembly._Assembler(V, GT.measure(Om, degree))


def relerr(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_noalloc_vector_matches_assemble_vector(GT, O):
    L = GT._lib
    print("case noalloc: Q2 hex (6,5,4) nodal f, then 2D Q1 (5,4) per-cell f", flush=True)
    cells1 = (6, 5, 4)
    m1, V1, a1 = make(GT, cells1, 2, 4, True, 0.15, 3)
    f1 = GT.inputs.normal(61, V1.nfree + V1.ndiri)
    p1 = O.Problem(list(cells1)).space(2, dirichlet=1)
    p1.set_coords(m1.coords)
    p1.quadrature(4)
    ref1 = {}
    for rows, n in ((L.GTK_FREE, V1.nfree), (L.GTK_DIRICHLET, V1.ndiri)):
        ref, _ = a1.ctx.assemble_vector(L.GTK_F_NODAL, f1, rows, n)
        out, st = a1.ctx.assemble_vector_noalloc(L.GTK_F_NODAL, f1, rows, n)
        assert np.array_equal(out, ref)
        assert st.nnz == n and st.kernel_ms > 0
        ref1[rows] = ref
    assert relerr(ref1[L.GTK_FREE], p1.assemble_vector(2, f1)) <= 1e-13

    # second call in a row: a smaller problem on another context
    cells2 = (5, 4)
    m2, V2, a2 = make(GT, cells2, 1, 2, False, 0.1, 4)
    f2 = GT.inputs.normal(62, m2.cells.shape[0])
    ref2, _ = a2.ctx.assemble_vector(L.GTK_F_CELL, f2, L.GTK_FREE, V2.nfree)
    out2, _ = a2.ctx.assemble_vector_noalloc(L.GTK_F_CELL, f2, L.GTK_FREE, V2.nfree)
    assert np.array_equal(out2, ref2)
    p2 = O.Problem(list(cells2)).space(1, dirichlet=0)
    p2.set_coords(m2.coords)
    p2.quadrature(2)
    assert relerr(ref2, p2.assemble_vector(1, f2)) <= 1e-13

    # the first problem again, with new load values
    f1b = GT.inputs.normal(63, V1.nfree + V1.ndiri)
    ref1b, _ = a1.ctx.assemble_vector(L.GTK_F_NODAL, f1b, L.GTK_FREE, V1.nfree)
    out1b, _ = a1.ctx.assemble_vector_noalloc(L.GTK_F_NODAL, f1b, L.GTK_FREE, V1.nfree)
    assert np.array_equal(out1b, ref1b)


def test_noalloc_boundary_load_clears_untouched_rows(GT):
    """A boundary load touches only boundary rows: the interior rows must be
    zero although the previous call left nonzero values in the scratch."""
    L = GT._lib
    print("case noalloc: 2D Q2 (6,5) boundary load after a volume load", flush=True)
    m, V, av = make(GT, (6, 5), 2, 4, False, 0.1, 5)
    _, _, ab = make(GT, (6, 5), 2, 4, False, 0.1, 5, boundary_measure=True)
    fv = GT.inputs.normal(71, V.nfree)
    ref_v, _ = av.ctx.assemble_vector(L.GTK_F_NODAL, fv, L.GTK_FREE, V.nfree)
    out_v, _ = av.ctx.assemble_vector_noalloc(L.GTK_F_NODAL, fv, L.GTK_FREE, V.nfree)
    assert np.array_equal(out_v, ref_v) and np.count_nonzero(out_v) == V.nfree
    ref_b, _ = ab.ctx.assemble_vector(L.GTK_F_CONST, np.array([2.0]), L.GTK_FREE, V.nfree)
    out_b, _ = ab.ctx.assemble_vector_noalloc(L.GTK_F_CONST, np.array([2.0]), L.GTK_FREE, V.nfree)
    assert np.array_equal(out_b, ref_b)
    assert 0 < np.count_nonzero(out_b) < V.nfree
    assert abs(out_b.sum() - 2.0 * 4.0) < 1e-12  # f = 2 over the perimeter 4 of the unit square


def test_noalloc_refuses_before_plan_and_beyond_scratch(GT):
    L = GT._lib
    print("case noalloc: call order and the 2048x2048 Q1 size limit", flush=True)
    _, V, a = make(GT, (4, 4), 1, 2, False, 0.0, 0)
    with pytest.raises(L.GtkError, match=r"\(-2\)"):
        a.ctx.assemble_vector_noalloc(L.GTK_F_CONST, np.array([1.0]), L.GTK_FREE, V.nfree)
    # 2049^2 rows (rounded up to 32) + as many nodal values > 8 Mi doubles;
    # a constant load (one value) fits
    _, Vb, ab = make(GT, (2048, 2048), 1, 2, False, 0.0, 0)
    n = Vb.nfree
    assert ((n + 31) // 32) * 32 + n > L.GTK_NOALLOC_MAX_DOUBLES >= ((n + 31) // 32) * 32 + 1
    ref, _ = ab.ctx.assemble_vector(L.GTK_F_CONST, np.array([1.0]), L.GTK_FREE, n)
    out, _ = ab.ctx.assemble_vector_noalloc(L.GTK_F_CONST, np.array([1.0]), L.GTK_FREE, n)
    assert np.array_equal(out, ref)
    assert abs(out.sum() - 1.0) < 1e-9
    with pytest.raises(L.GtkError, match=r"\(-1\).*scratch"):
        ab.ctx.assemble_vector_noalloc(L.GTK_F_NODAL, np.ones(n), L.GTK_FREE, n)
