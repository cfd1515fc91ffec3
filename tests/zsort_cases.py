"""z distributions for the z-order tests: uniform, one layer (all equal),
two interleaved layers, a layer with last-bit noise (distinct but nearly
equal keys), ascending and descending in the atom index, a few distinct
values.  All strictly inside (0, Lz) for Lz = 20."""
import numpy as np

LZ = 20.0

DISTS = ["uniform", "one_layer", "two_layers", "noisy_layer", "ascending", "descending",
         "few_values"]


def z_values(dist: str, n: int, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    if dist == "uniform":
        return rng.uniform(0.5, 19.5, n)
    if dist == "one_layer":
        return np.full(n, 5.0)
    if dist == "two_layers":
        return np.where(np.arange(n) % 2 == 0, 3.0, 7.0)
    if dist == "noisy_layer":
        # 5 + a few ulps either way: distinct keys, all within 1e-14
        ulp = np.spacing(5.0)
        return 5.0 + ulp * rng.integers(-4, 5, n)
    if dist == "ascending":
        return np.linspace(0.5, 19.5, n)
    if dist == "descending":
        return np.linspace(19.5, 0.5, n)
    if dist == "few_values":
        return rng.choice([0.75, 1.5, 2.0, 9.0, 9.0 + 1e-9, 12.25, 19.0], n)
    raise ValueError(dist)


def stable_order(z: np.ndarray) -> np.ndarray:
    """Ascending z, ties by index (the reference's sortperm)."""
    return np.lexsort((np.arange(z.shape[0]), z))
