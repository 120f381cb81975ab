"""Writes the small golden fixtures under tests/golden/ from the FP64 oracle.

The reference's own tests hold no fixtures or golden vectors (they compare against
analytic functions through quadrature), so these vectors are produced by the oracle
after it passes the analytic pins of tests/test_oracle_*.py.  They anchor the
oracle and the HIP path against regressions.

    python tests/golden/make_golden.py
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

from oracle import oracle as O  # noqa: E402


def main():
    out = {}
    for k in range(1, 5):
        out[f"dg_basis_k{k}"] = O.dg_basis(k)
    out["hier_dlf_k2_n3"] = O.hier_dlf(2, 3)
    out["hier_dlf_k3_n2"] = O.hier_dlf(3, 2)
    out["pos_dlf_k2_n2"] = O.pos_dlf(2, 2)
    out["hier2pos_k2_n2"] = O.hier2pos(2, 2)
    for d in (1, 2):
        A = O.d_matrix(2, d, 2, 3)
        out[f"d_matrix_D2_k2_n3_d{d}_colptr"] = A.indptr.astype(np.int64)
        out[f"d_matrix_D2_k2_n3_d{d}_rowval"] = A.indices.astype(np.int64)
        out[f"d_matrix_D2_k2_n3_d{d}_nzval"] = A.data
    # cfg1 shape (D=2 k=3 n=5): sin(2 pi x1) sin(2 pi x2), v0 = 0, 20 RK4 steps of dt = 0.1 * 2^-5
    D, k, n = 2, 3, 5
    s1 = O.proj1d(k, n, lambda x: np.sin(2 * np.pi * x))
    u0 = O.tensor_construct(D, k, n, [s1, s1])
    y0 = np.concatenate([u0, np.zeros_like(u0)])
    sysw = O.System("wave", D, k, n)
    out["wave_D2_k3_n5_y0"] = y0
    out["wave_D2_k3_n5_y20"] = sysw.rk4(y0, 0.1 * 2.0 ** -n, 20)
    # transport D=3 k=2 n=4, a = (1, 0.5, 0.25), 10 steps
    D, k, n, a = 3, 2, 4, [1.0, 0.5, 0.25]
    rng = np.random.default_rng(0)
    u0 = rng.standard_normal(O.get_size(D, k, n))
    syst = O.System("transport", D, k, n, a=a)
    out["transport_D3_k2_n4_u0"] = u0
    out["transport_D3_k2_n4_rhs"] = syst.rhs(u0)
    out["transport_D3_k2_n4_u10"] = syst.rk4(u0, 0.1 * 2.0 ** -n / 1.75, 10)
    np.savez_compressed(HERE / "oracle_golden.npz", **out)
    print(f"wrote {len(out)} arrays to {HERE / 'oracle_golden.npz'}")


if __name__ == "__main__":
    main()
