"""Multi-point constraints of the Transformation handler: the rigidDiaphragm
and equalDOF records of the case JSON and the per-DOF expansion tables the
analyses use.

``rigid_diaphragm`` writes the records in the shape Strut's harness gives them
(scripts/tcl_to_strut.py:2163-2212: constrained and retained DOFs and the
coefficient matrix from the nodal offsets); ``mpc_tables`` restates the
loader's table build (solver/run_case/loader.mojo:2230-2372): each DOF i
expands as u_i = sum_t coeff[t] u[dof_pool[t]] over t in
[row_offsets[i], row_offsets[i + 1]) -- the identity for an ordinary DOF, the
retained DOFs for a constrained one, nothing for a DOF fixed by a single-point
constraint.  The tables feed the device's collapse kernels
(``sgpu_mpc_*``); nothing here collapses a matrix.
"""

from __future__ import annotations

from typing import Dict, List, Sequence

import numpy as np


def rigid_diaphragm(coords: Dict[int, Sequence[float]], retained: int, constrained: Sequence[int], perp_dirn: int,
                    ndm: int = 3, ndf: int = 6) -> List[dict]:
    """rigidDiaphragm records (tcl_to_strut.py:2163-2250) of one retained
    node and its constrained nodes; coords maps node id -> (x, y, z)."""
    rx, ry, rz = (float(v) for v in (list(coords[retained]) + [0.0, 0.0])[:3])
    out = []
    for c in constrained:
        cx, cy, cz = (float(v) for v in (list(coords[c]) + [0.0, 0.0])[:3])
        dx, dy, dz = cx - rx, cy - ry, cz - rz
        if ndm == 3 and ndf == 6:
            if perp_dirn == 3:
                cd, rd, m = [1, 2, 6], [1, 2, 6], [[1.0, 0.0, -dy], [0.0, 1.0, dx], [0.0, 0.0, 1.0]]
            elif perp_dirn == 2:
                cd, rd, m = [1, 3, 5], [1, 3, 5], [[1.0, 0.0, dz], [0.0, 1.0, -dx], [0.0, 0.0, 1.0]]
            elif perp_dirn == 1:
                cd, rd, m = [2, 3, 4], [2, 3, 4], [[1.0, 0.0, -dz], [0.0, 1.0, dy], [0.0, 0.0, 1.0]]
            else:
                raise ValueError("rigidDiaphragm perpDirn must be 1, 2, or 3")
        elif ndm == 2 and ndf == 3:
            if perp_dirn == 3:
                cd, rd, m = [1, 2, 3], [1, 2, 3], [[1.0, 0.0, -dy], [0.0, 1.0, dx], [0.0, 0.0, 1.0]]
            elif perp_dirn in (1, 2):
                cd, rd, m = [perp_dirn], [perp_dirn], [[1.0]]
            else:
                raise ValueError("rigidDiaphragm perpDirn must be 1, 2, or 3")
        else:
            raise ValueError("rigidDiaphragm requires a 3D/6DOF or 2D/3DOF model")
        out.append({"type": "rigidDiaphragm", "retained_node": retained, "constrained_node": c,
                    "perp_dirn": perp_dirn, "constrained_dofs": cd, "retained_dofs": rd, "matrix": m})
    return out


def mpc_tables(n_nodes: int, ndf: int, sp_constrained: np.ndarray, mp_constraints: List[dict],
               node_index: Dict[int, int]):
    """The loader's expansion tables (loader.mojo:2230-2372).  sp_constrained:
    bool per DOF (single-point constraints); node_index: node id -> index.
    Returns (row_offsets, dof_pool, coeff_pool, slave, constrained) with
    constrained = SP-fixed or rigidDiaphragm-constrained DOFs."""
    total = n_nodes * ndf
    constrained = np.array(sp_constrained, dtype=bool).copy()
    row_count = np.ones(total, dtype=np.int64)
    dofs = [[i, -1, -1] for i in range(total)]
    coeffs = [[1.0, 0.0, 0.0] for _ in range(total)]
    slave = np.zeros(total, dtype=bool)
    controlled_nodes = {node_index[m["constrained_node"]] for m in mp_constraints}
    for m in mp_constraints:
        ri, ci = node_index[m["retained_node"]], node_index[m["constrained_node"]]
        if ri in controlled_nodes:
            raise ValueError(f"{m['type']} retained_node cannot also be constrained")
        if m["type"] == "equalDOF":
            for d in m["dofs"]:
                cdof, rdof = ci * ndf + d - 1, ri * ndf + d - 1
                if slave[cdof]:
                    raise ValueError("equalDOF constrained dof already controlled")
                row_count[cdof] = 1
                dofs[cdof] = [rdof, -1, -1]
                coeffs[cdof] = [1.0, 0.0, 0.0]
                slave[cdof] = True
        elif m["type"] == "rigidDiaphragm":
            for row, cd in enumerate(m["constrained_dofs"]):
                cdof = ci * ndf + cd - 1
                if slave[cdof]:
                    raise ValueError("rigidDiaphragm constrained dof already controlled")
                row_count[cdof] = len(m["retained_dofs"])
                dofs[cdof] = [-1, -1, -1]
                coeffs[cdof] = [1.0, 0.0, 0.0]
                for col, rd in enumerate(m["retained_dofs"]):
                    dofs[cdof][col] = ri * ndf + rd - 1
                    coeffs[cdof][col] = float(m["matrix"][row][col])
                slave[cdof] = True
                constrained[cdof] = True
        else:
            raise ValueError(f"unsupported mp constraint type: {m['type']}")
    row_offsets = np.zeros(total + 1, dtype=np.int32)
    pool = 0
    for i in range(total):
        row_offsets[i] = pool
        if constrained[i] and not slave[i]:
            continue
        pool += int(row_count[i])
    row_offsets[total] = pool
    dof_pool = np.full(pool, -1, dtype=np.int32)
    coeff_pool = np.zeros(pool)
    for i in range(total):
        if constrained[i] and not slave[i]:
            continue
        b = row_offsets[i]
        for t in range(int(row_count[i])):
            dof_pool[b + t] = dofs[i][t]
            coeff_pool[b + t] = coeffs[i][t]
    return row_offsets, dof_pool, coeff_pool, slave, constrained


def expansion_matrix(row_offsets, dof_pool, coeff_pool) -> np.ndarray:
    """T with u_full = T u_retained (T[i, dof_pool[t]] = coeff[t]); the
    collapse of a matrix K is T^T K T (used by the checks in tests/)."""
    n = len(row_offsets) - 1
    T = np.zeros((n, n))
    for i in range(n):
        for t in range(row_offsets[i], row_offsets[i + 1]):
            T[i, dof_pool[t]] += coeff_pool[t]
    return T


class TransformationHandler:
    """The Transformation handler's collapses on the device (sgpu_mpc_*),
    on a SectionRuntime's context."""

    def __init__(self, rt):
        import ctypes
        from ctypes import POINTER, c_double, c_int, c_int32, c_void_p

        self.rt = rt
        self.lib = rt.lib
        if not getattr(self.lib, "_mpc_declared", False):
            ip, dp = POINTER(c_int32), POINTER(c_double)
            sig = {
                "sgpu_mpc_set": (c_int, [c_void_p, c_int, ip, ip, dp, ip, ip]),
                "sgpu_mpc_collapse_matrix": (c_int, [c_void_p, dp, dp]),
                "sgpu_mpc_collapse_matrix_device": (c_int, [c_void_p, c_void_p, c_void_p]),
                "sgpu_mpc_collapse_vector": (c_int, [c_void_p, dp, dp]),
                "sgpu_mpc_enforce": (c_int, [c_void_p, dp]),
            }
            for name, (res, args) in sig.items():
                fn = getattr(self.lib, name)
                fn.restype = res
                fn.argtypes = args
            self.lib._mpc_declared = True
        self.n = 0

    def set(self, row_offsets, dof_pool, coeff_pool, constrained, slave) -> None:
        from ._native import dptr, iptr

        ro = np.ascontiguousarray(row_offsets, dtype=np.int32)
        dp = np.ascontiguousarray(dof_pool, dtype=np.int32)
        cp = np.ascontiguousarray(coeff_pool, dtype=np.float64)
        co = np.ascontiguousarray(constrained, dtype=np.int32)
        sl = np.ascontiguousarray(slave, dtype=np.int32)
        self.n = len(ro) - 1
        self.rt._check(self.lib.sgpu_mpc_set(self.rt.ctx, self.n, iptr(ro), iptr(dp), dptr(cp), iptr(co), iptr(sl)))

    def collapse_matrix(self, K) -> np.ndarray:
        from ._native import dptr

        K = np.ascontiguousarray(K, dtype=np.float64)
        out = np.zeros_like(K)
        self.rt._check(self.lib.sgpu_mpc_collapse_matrix(self.rt.ctx, dptr(K), dptr(out)))
        return out

    def collapse_matrix_device(self, d_K: int, d_out: int) -> None:
        self.rt._check(self.lib.sgpu_mpc_collapse_matrix_device(self.rt.ctx, d_K, d_out))

    def collapse_vector(self, F) -> np.ndarray:
        from ._native import dptr

        F = np.ascontiguousarray(F, dtype=np.float64)
        out = np.zeros_like(F)
        self.rt._check(self.lib.sgpu_mpc_collapse_vector(self.rt.ctx, dptr(F), dptr(out)))
        return out

    def enforce(self, u) -> np.ndarray:
        from ._native import dptr

        u = np.array(u, dtype=np.float64)
        self.rt._check(self.lib.sgpu_mpc_enforce(self.rt.ctx, dptr(u)))
        return u
