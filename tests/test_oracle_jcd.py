"""The oracle's restatement of the reference's Jacobi-Davidson with GMRES correction (oracle/jcd.py),
pinned by linear algebra: GMRES with enough steps and no early stop solves the system exactly
(numpy.linalg.solve), its early stop drops the step that met the tolerance (the reference's update
rule), and the eigensolver with and without GMRES converges to the lowest eigenvalue
(numpy.linalg.eigvalsh), orthogonal to given ortho vectors."""
import numpy as np

from oracle.jcd import gmres, jcd


def test_gmres_exact_and_early_stop():
    rng = np.random.default_rng(1)
    n = 6
    A = rng.normal(size=(n, n)) + n * np.eye(n)
    b = rng.normal(size=n)
    x0 = np.zeros(n)
    with np.errstate(divide="ignore", invalid="ignore"):
        x = gmres(lambda v: A @ v, b, x0, n + 1, 0.0)  # n Arnoldi steps, never stopped early
    assert np.allclose(x, np.linalg.solve(A, b), rtol=0, atol=1e-10)
    # a tolerance just below the start residual passes the first check and is met by the first
    # Arnoldi step; the update leaves that step out, so the start vector comes back unchanged
    tol = 0.999 * np.linalg.norm(b)
    y = gmres(lambda v: A @ v, b, x0, n + 1, tol)
    assert np.array_equal(y, x0)
    # with the tolerance met only at a later step, the update uses the steps before it
    z = gmres(lambda v: A @ v, b, x0, n + 1, 1e-3 * np.linalg.norm(b))
    assert not np.array_equal(z, x0) and np.linalg.norm(A @ z - b) < np.linalg.norm(b)


def test_jcd_lowest_eigenpair_with_and_without_gmres():
    rng = np.random.default_rng(2)
    n = 40
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    w = np.concatenate([[-5.0, -4.99, -4.98], np.linspace(-4.0, 3.0, n - 3)])
    A = (Q * w) @ Q.T
    v0 = rng.normal(size=n)
    th0, u0, it0, _ = jcd(A, v0, tol=1e-10, max_iter=40, gmres_steps=0)
    th3, u3, it3, _ = jcd(A, v0, tol=1e-10, max_iter=40, gmres_steps=4)
    for th, u in ((th0, u0), (th3, u3)):
        assert abs(th - w[0]) < 1e-9
        assert np.linalg.norm(A @ u - th * u) < 1e-8
    # the next state with the ground state projected out
    th1, u1, _, _ = jcd(A, v0, ortho=[u3], tol=1e-10, max_iter=40, gmres_steps=4)
    assert abs(th1 - w[1]) < 1e-8 and abs(u1 @ u3) < 1e-9
