"""cps_pressure_solve_csr_noalloc (static per-GPU buffers, no device
allocation) against cps_pressure_solve_csr on the same pressure systems:
two calls in a row, the second on a smaller system, must give the same
iteration counts and bit-identical solutions (both run the same kernels on
the same launch grids; the second call finds the first call's data in the
shared buffers).  Systems: the oracle's pressure matrices of density fields
after 3 SIMPLE steps (1024 x 512, the size limit's neighbourhood, then
96 x 48)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _system(pkg, orc, nx, ny, seed):
    cfg, gamma, inlet = pkg.configs.density_field(nx=nx, ny=ny, seed=seed, sigma=3.0)
    cfg.pressure_solver = pkg._types.CPS_PRESSURE_PCG
    ref = orc.OracleSimple(pkg._types, cfg, gamma, inlet, prec=64)
    ref.step(3)
    return ref.assemble("p")


def test_no_alloc_matches_allocating_entry_point(pkg, orc):
    for nx, ny, seed, maxit in ((1024, 512, 0, 200), (96, 48, 1, 2000)):
        A = _system(pkg, orc, nx, ny, seed)
        n = len(A["rhs"])
        tol = 1e-5 * float(np.linalg.norm(A["rhs"]))
        print(f"case noalloc n={n} nnz={A['row_ptr'][-1]} maxit={maxit}", flush=True)
        x_ref = np.zeros(n, np.float32)
        it_ref = pkg.pressure_solve_csr(A["row_ptr"], A["col_idx"], A["values"], A["rhs"], x_ref, tol, maxit)
        x = np.full(n, 0.0, np.float32)
        it = pkg.pressure_solve_csr_noalloc(A["row_ptr"], A["col_idx"], A["values"], A["rhs"], x, tol, maxit)
        print(f"  iterations {it} (allocating entry point {it_ref})", flush=True)
        assert it == it_ref
        assert np.array_equal(x, x_ref)
        assert np.isfinite(x).all() and np.abs(x).max() > 0


def test_no_alloc_rejects_sizes_beyond_its_limit(pkg):
    n = (1 << 19) + 1
    rp = np.arange(n + 1, dtype=np.int32)
    ci = np.arange(n, dtype=np.int32)
    val = np.ones(n, np.float32)
    b = np.ones(n, np.float32)
    x = np.zeros(n, np.float32)
    with pytest.raises(pkg._capi.CpsError) as e:
        pkg.pressure_solve_csr_noalloc(rp, ci, val, b, x, 1e-6, 10)
    assert e.value.code == pkg._types.CPS_ERR_ARG
