"""MI355X-native finite-element assembly for GalerkinToolkit-style problems.

Host-side mirror of the reference's interface for the assembly path
(GT.cartesian_mesh, GT.interior/boundary, GT.lagrange_space, GT.measure,
GT.assemble_matrix/assemble_vector, GT.update_matrix!) whose numeric work runs
in libgtk_hip.so (HIP kernels for gfx950 behind the C ABI of include/gtk_hip.h).
"""
from .assembly import (FP64, MIXED, DiscreteField, Elasticity, FaceConstantField, FormSum, InteriorPenalty, Laplace, Load,
                       Mass, Measure,
                       MatrixCache, Nitsche, NitscheLoad, SparseMatrixCSC, VectorCache, assemble_matrix, assemble_matrix_and_vector,
                       assemble_matrix_and_vector_with_free_and_dirichlet_columns,
                       assemble_matrix_with_free_and_dirichlet_columns, assemble_vector, face_constant_field,
                       linear_problem, measure, solve, update_matrix, update_matrix_and_vector,
                       update_matrix_and_vector_with_free_and_dirichlet_columns,
                       update_matrix_with_free_and_dirichlet_columns, update_vector)
from .mesh import Mesh, cartesian_mesh, perturb_interior_nodes
from .space import DIRICHLET, FREE, Domain, LagrangeSpace, boundary, interior, lagrange_space, skeleton
from .batch import MatrixBatch, assemble_matrices, autotune, autotune_matrix
from . import inputs, refcell

__all__ = [
    "FP64", "MIXED", "FREE", "DIRICHLET", "DiscreteField", "Elasticity", "FaceConstantField", "FormSum", "Laplace",
    "Load", "Nitsche", "NitscheLoad", "InteriorPenalty", "skeleton",
    "Mass", "Measure", "SparseMatrixCSC", "assemble_matrix", "assemble_matrix_with_free_and_dirichlet_columns",
    "assemble_vector", "face_constant_field", "linear_problem", "measure", "solve", "update_matrix", "Mesh",
    "MatrixCache", "VectorCache", "assemble_matrix_and_vector",
    "assemble_matrix_and_vector_with_free_and_dirichlet_columns", "update_matrix_and_vector",
    "update_matrix_and_vector_with_free_and_dirichlet_columns", "update_matrix_with_free_and_dirichlet_columns",
    "update_vector",
    "cartesian_mesh", "perturb_interior_nodes", "MatrixBatch", "assemble_matrices", "autotune", "autotune_matrix", "Domain",
    "LagrangeSpace", "boundary", "interior", "lagrange_space", "inputs", "refcell",
]
