"""The single-problem autotuner (batch.autotune_matrix) and its knob
gtk_set_launch: every launch shape it picks assembles the oracle's matrix
(FP64 within 1e-13, mixed within 1e-6), its choice is cached per element
type and reused, and shapes a kernel does not have are refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def relfro(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("mixed", 1e-6)])
def test_autotune_q2_hex(GT, O, tmp_path, precision, tol):
    print(f"case autotune Q2 hex {precision}", flush=True)
    cells = (6, 5, 6)
    m = GT.cartesian_mesh((0.0, 1.0) * 3, cells)
    m.coords = GT.perturb_interior_nodes(m, 0.15, 3)
    Om = GT.interior(m)
    V = GT.lagrange_space(Om, 2, dirichlet_boundary=GT.boundary(m))
    dO = GT.measure(Om, 4)
    kappa = GT.inputs.lognormal(5, m.cells.shape[0])
    form = GT.Laplace(dO, GT.face_constant_field(kappa, Om))
    prec = GT.FP64 if precision == "fp64" else GT.MIXED
    cache = str(tmp_path / "tune.json")
    r1 = GT.autotune_matrix(form, V, precision=prec, cache=cache)
    assert not r1["cached"] and len(r1["candidates_ms"]) >= 4
    assert r1["plan"]["matrix_cores"] == 1 and r1["plan"]["threads"] == 64 * r1["waves"]
    assert r1["plan"]["tile_cap"] == r1["max_cells"]
    p = O.Problem(list(cells)).space(2, dirichlet=1)
    p.set_coords(m.coords)
    p.quadrature(4)
    cp, rv, nz = p.assemble_matrix(O.LAPLACE, coef=kappa)
    A = GT.assemble_matrix(form, float, V, V, precision=prec)
    assert np.array_equal(A.colptr, cp) and np.array_equal(A.rowval, rv)
    assert relfro(A.nzval, nz) <= tol
    r2 = GT.autotune_matrix(form, V, precision=prec, cache=cache)
    assert r2["cached"] and (r2["waves"], r2["max_cells"]) == (r1["waves"], r1["max_cells"])
    # every candidate shape, applied directly, gives the same matrix
    a = A.device[0]
    code = GT._lib.GTK_LAPLACE
    lp = GT._lib.GTK_FP64 if precision == "fp64" else GT._lib.GTK_MIXED
    for w in (8, 12, 16):
        for cap in (0, 40):
            a.ctx.set_launch(code, lp, GT.FREE, GT.FREE, w, cap)
            vals, _ = a.ctx.assemble_matrix(code, lp, GT.FREE, GT.FREE, len(nz))
            assert relfro(vals, nz) <= tol, (w, cap)
    with pytest.raises(GT._lib.GtkError):
        a.ctx.set_launch(code, lp, GT.FREE, GT.FREE, 4, 0)  # matrix-core kernels: 8, 12 or 16 waves
    with pytest.raises(GT._lib.GtkError):
        a.ctx.set_launch(code, lp, GT.FREE, GT.FREE, 7, 0)


def test_autotune_valu_elasticity(GT, tmp_path):
    print("case autotune P2 tet elasticity", flush=True)
    m = GT.cartesian_mesh((0.0, 1.0) * 3, (3, 3, 3), simplexify=True)
    Om = GT.interior(m)
    V = GT.lagrange_space(Om, 2, tensor_size=(3,), dirichlet_boundary=GT.boundary(m, group_names=["2-face-5"]))
    dO = GT.measure(Om, 5)
    nc = m.cells.shape[0]
    lam = GT.face_constant_field(0.5 + GT.inputs.uniform(1, nc), Om)
    mu = GT.face_constant_field(0.5 + GT.inputs.uniform(2, nc), Om)
    form = GT.Elasticity(dO, lam, mu)
    ref = GT.assemble_matrix(form, float, V, V).nzval.copy()
    r = GT.autotune_matrix(form, V, cache=str(tmp_path / "t.json"))
    assert r["plan"]["matrix_cores"] == 0 and r["waves"] in (0, 4)
    A = GT.assemble_matrix(form, float, V, V)
    assert relfro(A.nzval, ref) <= 1e-13
