"""A regrid that flattens: the input shared by the golden generator
(tests/golden/make_golden.py, run against the compiled reference) and the
product's test.

After the initial data and the first regrid of a parity tree, every cell's
primitives are replaced by a smooth state: rho and eps within 1e-3 of 1,
velocities below 1e-3, a uniform field.  The refinement criterion of
amrtree::regrid (grid.cpp:744-795) then finds no density contrast above its
flatten limit (0.1) and no field gradient, so every leaf patch that is not
within r < 3 of the hole flags itself for erasure, and each parent whose
eight children are such leaves is flattened (ccrstr, grid.cpp:657-718:
volume-weighted restriction of U, con2prim, B_init), while the patches near
the hole refine once more.  U of the active zones is set to ptoc(W) first
(main.cpp:148-159) on both sides.
"""
import numpy as np

SEED = 424242
FIELD = (1e-3, -2e-3, 5e-4)


def flat_state(shape, seed=SEED) -> np.ndarray:
    """W [nb, 8, ncell] (rho, eps, v0..2, B0..2) for the flatten case."""
    rng = np.random.default_rng(seed)
    W = np.empty(shape)
    n = shape[0] * shape[2]
    W[:, 0] = 1.0 + 1e-3 * rng.random(n).reshape(shape[0], shape[2])
    W[:, 1] = 1.0 + 1e-3 * rng.random(n).reshape(shape[0], shape[2])
    for d in range(3):
        W[:, 2 + d] = 1e-3 * (rng.random(n).reshape(shape[0], shape[2]) - 0.5)
        W[:, 5 + d] = FIELD[d]
    return W
