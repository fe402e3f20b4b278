"""Inputs for the parity tests of the initial data's correctly rounded functions."""
import numpy as np


def ic_math_inputs(rng):
    """Inputs of the initial data's functions: polar and azimuthal angles,
    acos arguments (near +-1 too), pow bases and exponents, atan2 pairs."""
    th = np.concatenate([rng.uniform(0, np.pi, 6000), rng.uniform(-np.pi, np.pi, 6000),
                         np.array([np.pi / 2, np.pi, -np.pi, np.pi / 4, 1e-300, 0.0, -0.0])])
    c = np.concatenate([rng.uniform(-1, 1, 6000), 1 - np.ldexp(rng.uniform(0, 1, 2000), -rng.integers(1, 50, 2000)),
                        -1 + np.ldexp(rng.uniform(0, 1, 2000), -rng.integers(1, 50, 2000)),
                        np.array([0.0, -0.0, 1.0, -1.0, 0.5, -0.5])])
    px = np.concatenate([np.exp(rng.uniform(-30, 30, 6000)), 1 + rng.uniform(-1e-3, 1e-3, 2000)])
    py = np.concatenate([rng.choice([1.5, 2.5, 5 / 3, -0.5, 1 / (5 / 3 - 1), (5 / 3) / (5 / 3 - 1), 2.0, -1.0], 6000),
                         rng.uniform(-50, 50, 2000)])
    ay = np.concatenate([rng.uniform(-50, 50, 6000), np.ldexp(rng.uniform(-1, 1, 2000), -rng.integers(0, 60, 2000)),
                         np.array([0.0, -0.0, 0.0, -0.0, 1.0, 1e-16])])
    ax = np.concatenate([rng.uniform(-50, 50, 6000), rng.uniform(-50, 50, 2000),
                         np.array([1.0, 1.0, -1.0, -1.0, -0.0, 30.0])])
    return th, c, px, py, ay, ax
