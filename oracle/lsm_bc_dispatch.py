"""TEST INFRASTRUCTURE ONLY -- never part of the product path.

The face order of the reference's ghost-fill dispatch over a Grid
(src/boundary_conditions/lsm_boundary_conditions.c:19-886): which single-face
fills linearExtrapolationBC / signedLinearExtrapolationBC /
copyExtrapolationBC / homogeneousNeumannBC apply, in which order, for a
location index on a 2D or 3D grid.  Used by tests/ with the oracle's face
fills to check the HIP library's dispatch.

  3D (:129-300): 0..5 that face; 6, 7, 8 the lo then hi face of x, y, z;
                 9 the six faces x_lo, x_hi, y_lo, y_hi, z_lo, z_hi.
  2D (:25-127):  0..3 that face; 6, 7 the lo then hi face of x, y;
                 9 the four faces x_lo, x_hi, y_lo, y_hi.
  Anything else (and 4, 5, 8 in 2D): no face (the switch's empty default).
"""


def faces(num_dims, idx):
    if num_dims == 3:
        if 0 <= idx <= 5:
            return [idx]
        if 6 <= idx <= 8:
            return [2 * (idx - 6), 2 * (idx - 6) + 1]
        if idx == 9:
            return list(range(6))
        return []
    if num_dims == 2:
        if 0 <= idx <= 3:
            return [idx]
        if idx in (6, 7):
            return [2 * (idx - 6), 2 * (idx - 6) + 1]
        if idx == 9:
            return list(range(4))
        return []
    return []


# single-face Fortran symbol of each dispatch function, per dimension
FACE_SYMBOL = {
    "linearExtrapolationBC": "lsm%ddlinearextrapolation_",
    "signedLinearExtrapolationBC": "lsm%ddsignedlinearextrapolation_",
    "copyExtrapolationBC": "lsm%ddcopyextrapolation_",
    "homogeneousNeumannBC": "lsm%ddcopyextrapolation_",
}
