"""Test inputs: the package's seeded frame configurations."""
from proresraw_rs_amd.workloads import (  # noqa: F401
    config0, config1, seeded_curve, seeded_matrix, seeded_qmat)
