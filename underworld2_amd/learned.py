"""Learned coarse-grid correction (placeholder until a model is trained)."""


def load_default(stokes):
    return None
