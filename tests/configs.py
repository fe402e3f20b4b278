"""Parity configurations (reference-expressible Cartesian patch trees).

Each is (dom_l, dom_h, nqlt, regrid, a, Q): the tree of
``grid::amrtree(dom_l, dom_h, nqlt, M=1, a, Q, 5/3)`` with the initial data of
src/fluid/main.cpp:98-159 and an optional regrid (main.cpp:168).
"""
# pure floor atmosphere far from the hole, 8 patches (finite for ~3 steps)
R1 = ((-5.0, -5.0, 30.0), (5.0, 5.0, 40.0), 2, False, 0.9, 0.1)
# near-horizon box; one regrid refines every root: 72 patches, 64 leaves
R2 = ((2.0, -2.0, -2.0), (6.0, 2.0, 2.0), 2, True, 0.9, 0.1)
# centred box; regrid leaves two roots coarse: 227 patches on two levels
# (exercises coarse-to-fine ghost prolongation)
R3 = ((-10.0, -10.0, -10.0), (10.0, 10.0, 10.0), 3, True, 0.9, 0.1)
# torus box without regrid, 27 patches
R4 = ((0.0, -6.0, -6.0), (12.0, 6.0, 6.0), 3, False, 0.5, 0.0)
# single patch, for small fixtures
R0 = ((-5.0, -5.0, 30.0), (5.0, 5.0, 40.0), 1, False, 0.9, 0.1)
GAMMA = 5.0 / 3.0
CFL = 0.4
