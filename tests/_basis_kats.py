"""The reference's known-answer tests of the cosine and sine bases
(test/SpectralTransformer_test.py:64-310), restated as data: the bases, the
grid of each axis (end point included or not), the physical field and the
non-zero spectral coefficients it must transform to, on the 256 x 256 unit
box of test/conftest.py:12-21."""
import numpy as np

CE, COS, SIN = 0, 1, 2


def _grid(n_x, n_z, x_end, z_end):
    x = np.linspace(0, 1.0, n_x, endpoint=x_end)
    z = np.linspace(0, 1.0, n_z, endpoint=z_end)
    return np.meshgrid(x, z, indexing="ij")


# name: (bases, x end point, z end point, field(X, Z), {(row, col): value})
KATS = {
    # SpectralTransformer_test.py:64-92
    "cosine_x_periodic_z": ((COS, CE), True, False,
                            lambda X, Z: np.cos(np.pi * X) + 2.0 * np.cos(2 * np.pi * X),
                            {(1, 0): 1.0, (2, 0): 2.0, (-1, 0): 1.0, (-2, 0): 2.0}),
    # 95-121
    "periodic_x_cosine_z": ((CE, COS), False, True,
                            lambda X, Z: np.cos(np.pi * Z) + 2.0 * np.cos(2 * np.pi * Z),
                            {(0, 1): 1.0, (0, 2): 2.0}),
    # 124-163
    "cosine_both": ((COS, COS), True, True,
                    lambda X, Z: (np.cos(np.pi * Z) + 2.0 * np.cos(2 * np.pi * Z)
                                  + np.cos(np.pi * X)
                                  + 2.0 * np.cos(2 * np.pi * X) * np.cos(np.pi * Z)),
                    {(0, 1): 1.0, (0, 2): 2.0, (1, 0): 1.0, (2, 1): 2.0, (-1, 0): 1.0,
                     (-2, 1): 2.0}),
    # 166-197
    "sine_both": ((SIN, SIN), True, True,
                  lambda X, Z: (np.sin(np.pi * Z) * np.sin(np.pi * X)
                                + 2.0 * np.sin(2 * np.pi * Z) * np.sin(3 * np.pi * X)),
                  {(1, 1): 1.0, (-1, 1): -1.0, (3, 2): 2.0, (-3, 2): -2.0}),
    # 200-228
    "sine_x_periodic_z": ((SIN, CE), True, False,
                          lambda X, Z: np.sin(np.pi * X) + 2.0 * np.sin(2 * np.pi * X),
                          {(1, 0): 1.0, (-1, 0): -1.0, (2, 0): 2.0, (-2, 0): -2.0}),
    # 231-260
    "sine_x_cosine_z": ((SIN, COS), True, True,
                        lambda X, Z: np.sin(np.pi * X) + 2.0 * np.sin(2 * np.pi * X),
                        {(1, 0): 1.0, (-1, 0): -1.0, (2, 0): 2.0, (-2, 0): -2.0}),
    # 263-310
    "cosine_x_sine_z": ((COS, SIN), True, True,
                        lambda X, Z: (np.cos(np.pi * X) * np.sin(np.pi * Z)
                                      + 2.0 * np.cos(3 * np.pi * X) * np.sin(2 * np.pi * Z)
                                      + 3.0 * np.sin(2 * np.pi * Z)),
                        {(1, 1): 1.0, (-1, 1): 1.0, (0, 2): 3.0, (3, 2): 2.0, (-3, 2): 2.0}),
}


def case(name, n=256):
    """(bases, physical field, expected spectrum) of one known-answer test
    on an n x n grid (spectral shape (2nn+1, nm), nn = nm = (n-1)//3)."""
    bases, x_end, z_end, field, coeffs = KATS[name]
    X, Z = _grid(n, n, x_end, z_end)
    nn = nm = (n - 1) // 3
    spec = np.zeros((2 * nn + 1, nm), dtype=np.complex128)
    for (r, c), v in coeffs.items():
        spec[r, c] = v
    return bases, field(X, Z), spec
