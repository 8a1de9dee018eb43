"""Vector-valued (elasticity) drop-ins: elastic_tet10, elastic3d, modal3d."""
