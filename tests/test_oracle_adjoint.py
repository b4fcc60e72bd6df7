"""The oracle's record adjoint of the approximate form factor, pinned by central finite differences
of the oracle's forward (FP64) on a small mesh, complex DWBA branches, a rotation and q = 0."""
import numpy as np

K0 = 2 * np.pi / (1.23984 / 10.0)
AI = np.radians(0.2)


def _rec_from(A, c):
    s = np.linalg.norm(A, axis=1)
    rec = np.zeros((A.shape[0], 8))
    rec[:, 0] = s
    rec[:, 1:4] = A / s[:, None]
    rec[:, 4:7] = c
    return rec


def test_record_adjoint_matches_finite_differences(oracle_mod):
    O = oracle_mod
    from hipgisaxs_amd import meshes

    rng = np.random.default_rng(7)
    tri = meshes.convex_polyhedron(rng, np.array([6.0, 4.0, 3.0]), 40)
    rec = O.tri9_to_rec8(tri)
    A = rec[:, :1] * rec[:, 1:4]
    c = rec[:, 4:7].copy()
    g = O.qz_extended(O.qgrid_explicit(7, 6, -1.5, 0.0, 1.5, 1.5, AI, K0), K0, AI, 3e-6 + 2e-8j)
    g.qz[0] = 0.0  # one q = 0 point (F = V there)
    g.qx[:] = 0.0
    g.qy[0] = 0.0
    R = np.array([[0.8, -0.6, 0.0], [0.6, 0.8, 0.0], [0.0, 0.0, 1.0]])
    W = rng.normal(size=g.nq) + 1j * rng.normal(size=g.nq)

    def L(A_, c_):
        return float(np.real(np.sum(W * O.ff_approx(_rec_from(A_, c_), g, R))))

    grad = O.ff_approx_adjoint(rec, g, W, R)
    for t in (0, 7, 23, len(A) - 1):
        for k in range(6):
            A1, c1, A2, c2 = A.copy(), c.copy(), A.copy(), c.copy()
            base = A if k < 3 else c
            h = 1e-6 * max(1.0, abs(base[t, k % 3]))
            if k < 3:
                A1[t, k] += h
                A2[t, k] -= h
            else:
                c1[t, k - 3] += h
                c2[t, k - 3] -= h
            fd = (L(A1, c1) - L(A2, c2)) / (2 * h)
            assert abs(fd - grad[t, k]) <= 1e-6 * max(1.0, abs(fd)), (t, k, fd, grad[t, k])
