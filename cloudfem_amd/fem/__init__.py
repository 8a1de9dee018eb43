"""Scalar drop-ins: tet_poisson, tet_heat, tet10."""
