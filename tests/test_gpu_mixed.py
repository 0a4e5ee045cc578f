"""Mixed-precision PCG (FP32 inner PCG, FP64 reliable-update refinement).

Bars: converged flag with an FP64 true residual ||b - A x|| < tol ||r0||
recomputed independently here in FP64 on the host; the error against the
FP64 oracle solution is stated per case (printed) and bounded by what the
tolerance allows (relative error <= 1e3 * tol for these well-conditioned
cases); inner iteration count within 25% of the FP64 solver's with a
preconditioner, 50% without one (FP32 loses orthogonality faster there).

IC0 is tested on the graph with a 1e-3 shift only: on a Laplacian shifted by
1e-8 the FP32 rounding of the residual puts ~6e-8 relative noise into the
near-null constant mode, which an IC0 preconditioner (unlike Jacobi)
amplifies by ~1e8, and the FP32 inner process stagnates."""
import numpy as np
import pytest

from cases import PROBLEMS
from lap_amd import lib as L
from lap_amd import solver as S
from lap_amd.problems import grid_laplacian, rhs_normal

pytestmark = pytest.mark.gpu


def centred(x):
    return x - x.mean()


@pytest.mark.parametrize("kind,pname", [("line_jacobi", "grid3d_16"), ("none", "grid3d_16"),
                                        ("line_jacobi", "graph_big"), ("ichol", "graph_big"),
                                        ("none", "graph_big")])
def test_mixed_converges_to_fp64_tolerance(orc, kind, pname):
    print(f"[case] mixed kind={kind} problem={pname}", flush=True)
    Ah = PROBLEMS[pname]()
    b = rhs_normal(Ah.n, 31)
    A = S.DeviceCSR.from_csr(Ah)
    P = S.make_preconditioner(A, Ah, kind)
    tol = 1e-10
    m = S.pcg_mixed(A, S.to_dev(b), P, tol=tol, maxit=5000)
    d = S.pcg(A, S.to_dev(b), P, tol=tol, maxit=5000)
    x = m.x.cpu().numpy()
    true = np.linalg.norm(b - orc.matvec(Ah, x))
    r0 = np.linalg.norm(b)
    xo, ito, _, _ = orc.cg(Ah, b, kind, tol=tol, maxit=5000)
    err = np.linalg.norm(centred(x) - centred(xo)) / np.linalg.norm(centred(xo))
    print(f"[diag] mixed {kind}/{pname}: it {m.iterations} (fp64 {d.iterations}, oracle {ito}), "
          f"refinements {m.refinements}, true rel res {true / r0:.2e}, err vs FP64 oracle "
          f"{err:.2e}", flush=True)
    assert m.flag == 0
    assert true < tol * r0 * 1.01
    assert m.iterations <= (1.5 if kind == "none" else 1.25) * d.iterations + 5
    assert err <= 1e3 * tol


def test_mixed_256_tolerance_grid():
    """Shifted 3D Laplacian (FP32 would round its diagonal 6+1e-8 to 6 and
    make the matrix singular; the FP64 diagonal keeps it SPD)."""
    Ah = grid_laplacian([48, 48, 48], 1e-8)
    b = rhs_normal(Ah.n, 12345)
    A = S.DeviceCSR.from_csr(Ah)
    P = S.make_preconditioner(A, Ah, "line_jacobi")
    m = S.pcg_mixed(A, S.to_dev(b), P, tol=1e-10)
    d = S.pcg(A, S.to_dev(b), P, tol=1e-10)
    x, xd = m.x.cpu().numpy(), d.x.cpu().numpy()
    err = np.linalg.norm(centred(x) - centred(xd)) / np.linalg.norm(centred(xd))
    print(f"[diag] mixed 48^3: it {m.iterations} vs fp64 {d.iterations}, refinements "
          f"{m.refinements}, err vs FP64 {err:.2e}", flush=True)
    assert m.flag == 0 and m.true_res < 1e-10 * m.res0
    assert m.iterations <= 1.25 * d.iterations


def test_mixed_rejects_unsupported_preconditioner():
    Ah = grid_laplacian([10, 10], 1e-8)
    A = S.DeviceCSR.from_csr(Ah)
    P = S.make_preconditioner(A, Ah, "GS_std")
    with pytest.raises(L.LapError):
        S.pcg_mixed(A, S.to_dev(rhs_normal(Ah.n, 1)), P)
