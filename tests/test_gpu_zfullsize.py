"""The benchmark workloads at their full sizes, checked through properties
that hold at any size (the oracle is too slow there; the same kernels are
pinned against it on small meshes in test_gpu_parity.py):

- pattern sizes in closed form, nnz = prod((k^2 + 2k) n + 1) on an n^3 hex
  mesh of order k (test/assembly_tests.jl:72-73 counts the same way), and
  230 n^3 + 138 n^2 + 24 n + 1 per component pair for P2 tets;
- zero row sums of the Laplace matrix without Dirichlet rows (constants are
  in its kernel) and its symmetry;
- total mass and total load (f = 1) equal to the volume of the unit cube
  (test/problems_tests.jl:55-59), which interior perturbations keep;
- rigid-body modes (translations, a rotation) in the kernel of the
  elasticity matrix;
- the mixed form within 1e-6 of the FP64 form (relative Frobenius)."""
import numpy as np
import pytest

from helpers import dof_coordinates

pytestmark = pytest.mark.gpu
SEED = 20261202


def relfro(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def hex_mesh(GT, n, perturb):
    m = GT.cartesian_mesh((0.0, 1.0) * 3, (n, n, n))
    if perturb:
        m.coords = GT.perturb_interior_nodes(m, perturb, SEED)
    return m


def laplace_properties(A):
    S = A.to_scipy()
    scale = np.abs(A.nzval).max()
    assert np.abs(S @ np.ones(S.shape[1])).max() <= 1e-12 * scale
    D = (S - S.T).tocsc()
    assert (np.abs(D.data).max() if D.nnz else 0.0) <= 1e-12 * scale


def test_q2_hex_64_perturbed(GT):
    print("case fullsize q2 hex 64^3 perturbed", flush=True)
    n = 64
    m = hex_mesh(GT, n, 0.2)
    Om = GT.interior(m)
    V = GT.lagrange_space(Om, 2)
    dO = GT.measure(Om, 4)
    A = GT.assemble_matrix(GT.Laplace(dO), float, V, V, precision=GT.FP64)
    assert A.nzval.size == (8 * n + 1) ** 3
    laplace_properties(A)
    Am = GT.assemble_matrix(GT.Laplace(dO), float, V, V, precision=GT.MIXED)
    assert np.array_equal(Am.rowval, A.rowval) and np.array_equal(Am.colptr, A.colptr)
    assert relfro(Am.nzval, A.nzval) <= 1e-6
    M = GT.assemble_matrix(GT.Mass(dO), float, V, V, precision=GT.FP64)
    assert abs(np.sum(M.nzval) - 1.0) <= 1e-12
    b = GT.assemble_vector(GT.Load(dO, 1.0), float, V)
    assert abs(np.sum(b) - 1.0) <= 1e-12


def test_q1_hex_128_lognormal(GT):
    print("case fullsize q1 hex 128^3 lognormal kappa", flush=True)
    n = 128
    m = hex_mesh(GT, n, 0.0)
    Om = GT.interior(m)
    V = GT.lagrange_space(Om, 1)
    dO = GT.measure(Om, 2)
    kappa = GT.inputs.lognormal(SEED, m.cells.shape[0])
    form = GT.Laplace(dO, GT.face_constant_field(kappa, Om))
    A = GT.assemble_matrix(form, float, V, V, precision=GT.FP64)
    assert A.nzval.size == (3 * n + 1) ** 3
    laplace_properties(A)
    Am = GT.assemble_matrix(form, float, V, V, precision=GT.MIXED)
    assert relfro(Am.nzval, A.nzval) <= 1e-6


def test_p2_tet_elasticity_48(GT):
    print("case fullsize p2 tet elasticity 48^3 x 6", flush=True)
    n = 48
    m = GT.cartesian_mesh((0.0, 1.0) * 3, (n, n, n), simplexify=True)
    Om = GT.interior(m)
    V = GT.lagrange_space(Om, 2, tensor_size=(3,))
    dO = GT.measure(Om, 5)
    nc = m.cells.shape[0]
    lam = 0.5 + 1.5 * GT.inputs.uniform(SEED, nc)
    mu = 0.5 + 1.5 * GT.inputs.uniform(SEED + 1, nc)
    form = GT.Elasticity(dO, GT.face_constant_field(lam, Om), GT.face_constant_field(mu, Om))
    K = GT.assemble_matrix(form, float, V, V, precision=GT.FP64)
    # scalar P2 tets: nnz = 230 n^3 + 138 n^2 + 24 n + 1 (the oracle gives it
    # for n = 2..5); three components coupled pairwise: 9 times that
    assert K.nzval.size == 9 * (230 * n ** 3 + 138 * n ** 2 + 24 * n + 1)
    S = K.to_scipy()
    scale = np.abs(K.nzval).max()
    xf, _ = dof_coordinates(GT, m, V)
    nodes = xf[0::3]  # the three components of a node are consecutive
    modes = []
    for d in range(3):
        r = np.zeros(S.shape[1])
        r[d::3] = 1.0
        modes.append(r)
    rot = np.zeros(S.shape[1])  # rotation about the z axis: (-y, x, 0)
    rot[0::3] = -nodes[:, 1]
    rot[1::3] = nodes[:, 0]
    modes.append(rot)
    for r in modes:
        assert np.abs(S @ r).max() <= 1e-11 * scale * max(1.0, np.abs(r).max())
    D = (S - S.T).tocsc()
    assert (np.abs(D.data).max() if D.nnz else 0.0) <= 1e-12 * scale
    Km = GT.assemble_matrix(form, float, V, V, precision=GT.MIXED)
    assert relfro(Km.nzval, K.nzval) <= 1e-6
