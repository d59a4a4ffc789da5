"""Cubature configuration (mirrors bayesnewton/cubature.py:30-37).

Only Gauss-Hermite is on the hot path; the nodes are computed inside
libbn_hip.so (Golub-Welsch, bn_gauss_hermite) and only the count travels.
"""


class GaussHermite:
    """cubature.py:30-37 (default 20 points)."""

    def __init__(self, dim=None, num_cub_points=20):
        if not 1 <= int(num_cub_points) <= 64:
            raise ValueError("num_cub_points must be in [1, 64]")
        self.num_cub_points = int(num_cub_points)


def num_points(cubature):
    if cubature is None:
        return 20
    if isinstance(cubature, GaussHermite):
        return cubature.num_cub_points
    raise NotImplementedError("only Gauss-Hermite cubature is implemented on the HIP path")
