"""Boundary conditions, mirroring ``underworld.conditions`` (PY/conditions/_conditions.py:63-107).

``DirichletCondition(variable, indexSetsPerDof)`` fixes, for each dof i of ``variable``,
the nodes of ``indexSetsPerDof[i]`` to the values currently held in ``variable.data``
(free slip = the normal component fixed to zero on each wall).
"""
import numpy as np


class DirichletCondition:
    def __init__(self, variable, indexSetsPerDof):
        if not isinstance(indexSetsPerDof, (tuple, list)):
            indexSetsPerDof = (indexSetsPerDof,)
        if len(indexSetsPerDof) != variable.nodeDofCount:
            raise ValueError("indexSetsPerDof must provide one entry (or None) per dof of the variable")
        self.variable = variable
        self.indexSetsPerDof = tuple(indexSetsPerDof)

    def mask(self, nnodes):
        """uint8 per node, bit i set where dof i is constrained."""
        m = np.zeros(nnodes, dtype=np.uint8)
        for d, s in enumerate(self.indexSetsPerDof):
            if s is None:
                continue
            idx = s.data if hasattr(s, "data") else np.asarray(s, dtype=np.int64)
            m[idx] |= np.uint8(1 << d)
        return m


class NeumannCondition:
    def __init__(self, *args, **kwargs):
        raise NotImplementedError("NeumannCondition (surface tractions) is not provided by this engine")
