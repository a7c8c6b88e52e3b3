from __future__ import annotations

import numpy as np
from numpy.typing import NDArray

class KDTree:
    data: NDArray[np.float64]
    ndim: int
    n_points: int
    leafsize: int

    def __init__(self, data: np.ndarray, *, leafsize: int = 32, parallel: bool = False) -> None: ...
    def __len__(self) -> int: ...
    def __repr__(self) -> str: ...
    def query(
        self,
        x: np.ndarray,
        *,
        k: int = 1,
        p: float = 2.0,
        max_distance: float | None = None,
        eps: float = 0.0,
        parallel: bool = False,
        fp32_screen: bool = False,
    ) -> tuple[NDArray[np.float64], NDArray[np.int64]]: ...
