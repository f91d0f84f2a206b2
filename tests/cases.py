"""Test operators shared by the CPU pins and the GPU parity tests."""
import numpy as np

import gdmr


def shift_plus_diag(n, alpha, seed=0):
    """Cyclic shift plus a jittered diagonal alpha (1 + U(-0.5, 0.5)), CSR with sorted columns:
    restarted GMRES nearly stalls on it (per-cycle ratios above 0.995), so the stagnation restart
    of arnoldiSetupDMR (gDMR.cu:2586-2632) fires."""
    d = alpha * (1.0 + gdmr.rand_uniform(seed, 0, n, -0.5, 0.5))
    cols, vals = [], []
    for i in range(n):
        for j, v in sorted([(i, d[i]), ((i + 1) % n, 1.0)]):
            cols.append(j)
            vals.append(v)
    rp = np.arange(0, 2 * n + 1, 2, dtype=np.int32)
    return gdmr.CsrMatrix(n, rp, np.array(cols, np.int32), np.array(vals))


# (n, alpha, m, k, matRank): both stall into several stagnation restarts with k > 0.
STAG_CASES = [(64, 0.2, 12, 4, 40), (128, 0.2, 20, 6, 60)]
