"""Skybox cube maps for the wormhole scene.

The reference loads two PNG skies (main/Wormhole.source/source_model.jl:10-20:
"NASA_Sky_4096.png" on the near side, "Red_Nebula.png" on the far side) from a
texture folder that it does not ship (README.md:51).  A skybox here is the same
horizontal-cross cube map, 8-bit RGB, 3N rows x 4N columns; `synthetic_cubemap`
makes a deterministic stand-in of any size (a smooth coloured sky plus a star
field) for tests and benchmarks.  `cross_faces` gives the pixel rectangles of the
six faces in the layout of cubemap_coord_calc (main/gen_geometry.jl:347-370).
"""
import numpy as np


def cross_faces(N):
    """{face: (row0, col0)} of the N x N blocks, rows counted from the top of the image
    (the reference indexes row 3N + 1 - j)."""
    # absolute (i, j) offsets of cubemap_coord_calc, j counted from the bottom
    off = {1: (0, 1), 2: (1, 1), 3: (2, 1), 4: (3, 1), 5: (1, 0), 6: (1, 2)}
    return {f: ((2 - dj) * N, di * N) for f, (di, dj) in off.items()}


def synthetic_cubemap(N, seed=0, tint=(1.0, 1.0, 1.0), star_fraction=0.004):
    """(3N, 4N, 3) uint8 cube map: a smooth two-colour gradient over each face, seeded
    stars, zero outside the cross."""
    rng = np.random.default_rng(seed)
    img = np.zeros((3 * N, 4 * N, 3), dtype=np.float64)
    y, x = np.mgrid[0:N, 0:N] / max(1, N - 1)
    tint = np.asarray(tint, dtype=np.float64)
    for f, (r0, c0) in cross_faces(N).items():
        ph = rng.uniform(0, 2 * np.pi, 3)
        base = np.stack([0.5 + 0.5 * np.sin(2 * np.pi * (x * (f % 3 + 1)) + ph[0]) * np.cos(np.pi * y + ph[1]),
                         0.5 + 0.5 * np.sin(2 * np.pi * (y * (f % 2 + 1)) + ph[2]),
                         0.5 + 0.5 * np.cos(np.pi * (x + y) * f)], axis=-1)
        block = 40.0 * base * tint
        stars = rng.random((N, N)) < star_fraction
        block[stars] = rng.uniform(150, 255, (int(stars.sum()), 1)) * np.array([1.0, 0.95, 0.9])
        img[r0:r0 + N, c0:c0 + N] = block
    return np.clip(np.rint(img), 0, 255).astype(np.uint8)


def default_skyboxes(N=1024):
    """Near side (blue-white sky) and far side (red nebula-like) stand-ins."""
    return (synthetic_cubemap(N, seed=1, tint=(0.8, 0.9, 1.2)),
            synthetic_cubemap(N, seed=2, tint=(1.6, 0.5, 0.4), star_fraction=0.002))
