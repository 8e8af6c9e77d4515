"""Device symbolic phase: the CSC pattern built on the GPU equals the host
symbolic phase (pinned bit for bit against a COO assembly in
tests/test_plan_host.py and against the oracle in tests/test_gpu_parity.py),
assembling on it gives the same values, and the device cell-local-entry ->
nonzero map points every local entry at its (row, column) in the block."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    # cells, order, simplex, ncomp, dirichlet
    ((7, 5), 1, False, 1, True),
    ((6, 4), 2, False, 1, False),
    ((5, 6), 3, False, 1, True),
    ((4, 3, 5), 1, False, 1, True),
    ((3, 4, 3), 2, False, 1, True),
    ((3, 3, 2), 2, True, 1, False),
    ((2, 3, 2), 2, True, 3, True),
    ((6, 5), 2, True, 2, True),
]


def make(GT, cells, order, simplex, ncomp, diri):
    m = GT.cartesian_mesh(tuple([0.0, 1.0] * len(cells)), cells, simplexify=simplex)
    Om = GT.interior(m)
    V = GT.lagrange_space(Om, order, tensor_size=(ncomp,) if ncomp > 1 else None,
                          dirichlet_boundary=GT.boundary(m) if diri else None)
    return m, Om, V


def contexts(GT, m, Om, V, degree):
    from gtk_amd import assembly as AS
    a_host = AS._Assembler(V, GT.measure(Om, degree))
    a_dev = AS._Assembler(V, GT.measure(Om, degree))
    return a_host, a_dev


@pytest.mark.parametrize("cells,order,simplex,ncomp,diri", CASES)
def test_device_pattern_equals_host(GT, cells, order, simplex, ncomp, diri):
    print(f"case symbolic {cells} order={order} simplex={simplex} ncomp={ncomp} diri={diri}", flush=True)
    m, Om, V = make(GT, cells, order, simplex, ncomp, diri)
    a_host, a_dev = contexts(GT, m, Om, V, 2 * order)
    blocks = [(1, 1), (1, 2)] if diri else [(1, 1)]
    for rows, cols in blocks:
        nnz_h, nc_h = a_host.ctx.build_pattern(rows, cols)
        nnz_d, nc_d, ms = a_dev.ctx.build_pattern_device(rows, cols)
        assert (nnz_h, nc_h) == (nnz_d, nc_d)
        cp_h, rv_h = a_host.ctx.get_pattern(rows, cols, nnz_h, nc_h)
        cp_d, rv_d = a_dev.ctx.get_pattern(rows, cols, nnz_d, nc_d)
        assert np.array_equal(cp_h, cp_d) and np.array_equal(rv_h, rv_d)
        # the local map: every in-block entry points at its row within its column
        nd = V.cell_dofs.shape[1]
        M = a_dev.ctx.local_nnz_map(rows, cols, m.cells.shape[0], nd)
        cd = V.cell_dofs
        rsel = cd > 0 if rows == 1 else cd < 0
        csel = cd > 0 if cols == 1 else cd < 0
        inblock = rsel[:, None, :] & csel[:, :, None]  # [cell, b, a]
        assert np.array_equal(M > 0, inblock)
        idx = M[inblock] - 1
        r = np.broadcast_to(np.abs(cd)[:, None, :], M.shape)[inblock]
        c = np.broadcast_to(np.abs(cd)[:, :, None], M.shape)[inblock]
        assert np.array_equal(rv_d[idx], r)
        assert np.all(cp_d[c - 1] - 1 <= idx) and np.all(idx < cp_d[c] - 1)


def test_assembly_on_device_pattern(GT):
    """Values assembled on the device-built pattern equal those on the host
    pattern (same plan, same kernel)."""
    print("case symbolic assembly", flush=True)
    m, Om, V = make(GT, (4, 4, 4), 2, False, 1, True)
    m.coords = GT.perturb_interior_nodes(m, 0.2, 3)
    a_host, a_dev = contexts(GT, m, Om, V, 4)
    out = []
    for a, dev in ((a_host, False), (a_dev, True)):
        if dev:
            nnz, _, _ = a.ctx.build_pattern_device(1, 1)
        else:
            nnz, _ = a.ctx.build_pattern(1, 1)
        for prec in (0, 1):
            v, _ = a.ctx.assemble_matrix(0, prec, 1, 1, nnz)
            out.append(v)
    assert np.array_equal(out[0], out[2]) and np.array_equal(out[1], out[3])


def test_device_pattern_timing(GT):
    """Q2 hexahedra 40^3: both phases timed (reported, not asserted)."""
    m, Om, V = make(GT, (40, 40, 40), 2, False, 1, True)
    a_host, a_dev = contexts(GT, m, Om, V, 4)
    t0 = time.perf_counter()
    nnz_h, _ = a_host.ctx.build_pattern(1, 1)
    host_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    nnz_d, _, dev_ms = a_dev.ctx.build_pattern_device(1, 1)
    call_ms = (time.perf_counter() - t0) * 1e3
    assert nnz_h == nnz_d
    print(f"Q2 hex 40^3 pattern: host {host_ms:.1f} ms, device kernels {dev_ms:.2f} ms "
          f"(call with host copy {call_ms:.1f} ms), nnz {nnz_d}", flush=True)
