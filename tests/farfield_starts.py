"""Cone start points for the far-field tests (fragment-like points of a scene)."""
import numpy as np


def starts(cv, scene, n, seed):
    import importlib
    ff = importlib.import_module("cloud_renderer_amd.farfield")
    return ff.fragment_starts(scene, n, np.random.Generator(np.random.PCG64(seed)))
