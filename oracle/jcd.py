"""Restatement of the reference's Jacobi-Davidson eigensolver with its GMRES correction, on a dense
matrix (test infrastructure: the checker of mq_davidson's iteration, never the product).

Follows (paths relative to dmrg/framework/):
  * solve_ietl_jcd (dmrg/solver/ietl_jacobi_davidson.h:35-76): Gram-Schmidt of the ortho vectors,
    basic_iteration(max_iter, tol, tol);
  * jacobi_davidson::calculate_eigenvalue (ietl/jacobi.h:386-474): modified Gram-Schmidt with one
    refinement at kappa = 0.25, projection onto the complement of the ortho vectors, the Rayleigh
    matrix column, its lowest eigenpair (dsyevx in get_extremal_eigenvalue; numpy's eigh here),
    u = V s, r = P(V^A s) - theta u, the stop rule of basic_iteration::finished
    (ietl/iteration.h:44-57), the correction with rel_tol = 2^-(iterations + 1);
  * jcd_gmres_solver (ietl/jacobi.h:126-164): t0 = -r + (r.u / u.u) u, then with max_iter > 0
    t = ietl_gmres(op, -r, t0, rel_tol), op = (1 - uu*)(A - theta)(1 - uu*)
    (jcd_solver_operator, jacobi.h:263-281);
  * ietl_gmres (ietl/gmres.h:40-145): Arnoldi with modified Gram-Schmidt, Givens rotations, at most
    max_iter - 1 steps, the solution updated from the columns before the step that meets the
    tolerance.
"""
import numpy as np


def _rot(dx, dy, cs, sn):
    return cs * dx + sn * dy, -sn * dx + cs * dy


def _gen_rot(dx, dy):
    if dy == 0.0:
        return 1.0, 0.0
    if abs(dy) > abs(dx):
        t = dx / dy
        sn = 1.0 / np.sqrt(1.0 + t * t)
        return t * sn, sn
    t = dy / dx
    cs = 1.0 / np.sqrt(1.0 + t * t)
    return cs, t * cs


def gmres(op, b, x0, max_iter, abs_tol):
    """ietl_gmres::operator() (gmres.h:89-145)"""
    m = max_iter
    s = np.zeros(m + 1)
    cs = np.zeros(m + 1)
    sn = np.zeros(m + 1)
    H = np.zeros((m + 1, m + 1))
    v = [None] * (m + 1)
    r = b - op(x0)
    s[0] = np.linalg.norm(r)
    if abs(s[0]) < abs_tol:
        return x0
    v[0] = r / s[0]
    i = 0
    while i < m - 1:
        w = op(v[i])
        for k in range(i + 1):
            H[k, i] = w @ v[k]
            w = w - H[k, i] * v[k]
        H[i + 1, i] = np.linalg.norm(w)
        v[i + 1] = w / H[i + 1, i]
        for k in range(i):
            H[k, i], H[k + 1, i] = _rot(H[k, i], H[k + 1, i], cs[k], sn[k])
        cs[i], sn[i] = _gen_rot(H[i, i], H[i + 1, i])
        H[i, i], H[i + 1, i] = _rot(H[i, i], H[i + 1, i], cs[i], sn[i])
        s[i], s[i + 1] = _rot(s[i], s[i + 1], cs[i], sn[i])
        if abs(s[i + 1]) < abs_tol:
            break
        i += 1
    y = s[:i].copy()
    for a in range(i - 1, -1, -1):
        y[a] /= H[a, a]
        for c in range(a - 1, -1, -1):
            y[c] -= H[c, a] * y[a]
    x = x0.copy()
    for k in range(i):
        x = x + y[k] * v[k]
    return x


def jcd(A, v0, ortho=(), tol=1e-8, max_iter=8, gmres_steps=0):
    """lowest eigenpair of the symmetric matrix A from v0: (theta, u, iterations, thetas per
    iteration)"""
    n = A.shape[0]
    O = [np.array(o, float) for o in ortho][: max(0, n - 1)]
    for a in range(1, len(O)):
        for b in range(a):
            O[a] = O[a] - (O[b] @ O[a]) / (O[b] @ O[b]) * O[b]

    def project(x):
        for o in O:
            x = x - (o @ x) / (o @ o) * o
        return x

    V = [None] * max_iter
    VA = [None] * max_iter
    M = np.zeros((max_iter, max_iter))
    V[0] = project(np.array(v0, float))
    it = 0
    thetas = []
    while True:
        t = V[it]
        tau = np.linalg.norm(t)
        for i in range(it):
            t = t - (V[i] @ t) * V[i]
        if np.linalg.norm(t) < 0.25 * tau:
            for i in range(it):
                t = t - (V[i] @ t) * V[i]
        t = project(t)
        t = t / np.linalg.norm(t)
        V[it] = t
        VA[it] = A @ t
        for i in range(it + 1):
            M[i, it] = V[i] @ VA[it]
        dim = it + 1
        Ms = np.triu(M[:dim, :dim]) + np.triu(M[:dim, :dim], 1).T
        w, S = np.linalg.eigh(Ms)
        theta, s = w[0], S[:, 0]
        u = sum(V[j] * s[j] for j in range(dim))
        uA = project(sum(VA[j] * s[j] for j in range(dim)))
        r = uA - theta * u
        it += 1
        thetas.append(theta)
        rn = np.linalg.norm(r)
        if rn <= tol * abs(theta) or rn < tol or it >= max_iter:
            return theta, u, it, thetas
        rel_tol = 1.0 / 2.0 ** (it + 1)
        tc = -r + (r @ u) / (u @ u) * u
        if gmres_steps > 0:
            def op(x):
                t2 = x - (u @ x) * u
                y = A @ t2 - theta * t2
                return y - (u @ y) * u
            tc = gmres(op, -r, tc, gmres_steps, rel_tol)
        V[it] = tc
