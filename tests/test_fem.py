"""Element matrices of the model generators, checked against mechanics.

Each element must have exactly its rigid-body null space, reproduce the
strain energy of a uniform strain field exactly (patch test: a wrong sign or
a dropped term in B or D changes the energy), and, for the beam, the closed-
form stiffnesses of a cantilever.
"""
import numpy as np

from sprgpu import fem


def _rank(K, rel=1e-10):
    w = np.linalg.eigvalsh(K)
    return int(np.sum(w > rel * w.max())), w


def test_hex8_rigid_modes_and_rank():
    K = fem.hex8_stiffness(1.0, 0.3)
    assert np.allclose(K, K.T)
    r, w = _rank(K)
    assert r == 18 and w.min() > -1e-12
    X = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                  [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    for t in np.eye(3):
        u = np.tile(t, 8)
        assert np.allclose(K @ u, 0, atol=1e-12)
    for axis in np.eye(3):
        u = np.cross(axis, X).ravel()
        assert np.allclose(K @ u, 0, atol=1e-12)


def test_hex8_uniform_strain_energy():
    E, nu = 2.0, 0.3
    K = fem.hex8_stiffness(E, nu)
    X = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                  [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    G = np.random.default_rng(0).normal(size=(3, 3))
    eps = 0.5 * (G + G.T)                      # symmetric strain tensor
    u = (X @ eps.T).ravel()                    # u = eps x
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    energy = 0.5 * (lam * np.trace(eps) ** 2 + 2 * mu * np.sum(eps * eps))  # unit volume
    assert np.isclose(0.5 * u @ K @ u, energy, rtol=1e-12)


def test_quad4_rigid_modes_and_patch():
    E, nu, t = 1.5, 0.25, 1.0
    K = fem.quad4_stiffness(E, nu, t)
    r, _ = _rank(K)
    assert r == 5
    X = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], float)
    assert np.allclose(K @ np.tile([1, 0], 4), 0)
    assert np.allclose(K @ np.tile([0, 1], 4), 0)
    assert np.allclose(K @ np.column_stack([-X[:, 1], X[:, 0]]).ravel(), 0)
    eps = np.array([[0.3, 0.1], [0.1, -0.2]])
    u = (X @ eps.T).ravel()
    c = E / (1 - nu * nu)
    sxx = c * (eps[0, 0] + nu * eps[1, 1])
    syy = c * (eps[1, 1] + nu * eps[0, 0])
    sxy = c * (1 - nu) / 2 * 2 * eps[0, 1]
    energy = 0.5 * (sxx * eps[0, 0] + syy * eps[1, 1] + sxy * 2 * eps[0, 1]) * t
    assert np.isclose(0.5 * u @ K @ u, energy, rtol=1e-12)


def test_beam_closed_forms():
    EA, EI, GJ, L = 3.0, 2.0, 1.5, 1.7
    d = np.array([0.6, -0.48, 0.64])
    d /= np.linalg.norm(d)
    K = fem.beam3d_stiffness(np.zeros(3), L * d, EA, EI, GJ)
    r, _ = _rank(K)
    assert r == 6
    # cantilever clamped at node 0: tip force P along the axis -> P L / EA;
    # transverse -> P L^3 / (3 EI); axial torque -> T L / GJ
    Kff = K[6:, 6:]
    perp = np.cross(d, [0, 0, 1.0])
    perp /= np.linalg.norm(perp)
    u_ax = np.linalg.solve(Kff, np.concatenate([d, np.zeros(3)]))
    assert np.isclose(u_ax[:3] @ d, L / EA, rtol=1e-10)
    u_tr = np.linalg.solve(Kff, np.concatenate([perp, np.zeros(3)]))
    assert np.isclose(u_tr[:3] @ perp, L ** 3 / (3 * EI), rtol=1e-10)
    u_t = np.linalg.solve(Kff, np.concatenate([np.zeros(3), d]))
    assert np.isclose(u_t[3:] @ d, L / GJ, rtol=1e-10)


def test_hex8_grid_counts():
    m = fem.hex8_grid(20)
    assert m.nanod == 21 ** 3 and m.nel == 8000
    assert int(np.sum(m.msc)) == 26460          # face z=0 clamped
    m40 = fem.hex8_grid(40, e_sigma=0.5, e_seed=1)
    assert int(np.sum(m40.msc)) == 201720
    assert np.isclose(np.log(m40.escale).std(), 0.5, atol=0.01)


def test_quad4_grid_counts():
    m = fem.quad4_grid(1000)
    assert int(np.sum(m.msc)) == 2002000


def test_hex8_stiffness_at_pins_the_element_formula():
    """The general (isoparametric) hex8 element against what the mathematics
    fixes: on the unit cube it is hex8_stiffness; on a distorted element it
    is symmetric positive semi-definite with exactly the six rigid-body modes
    as its null space; on a parallelepiped (an affine image of the cube) a
    linear displacement field u = G x gives the exact strain energy
    vol * eps:D:eps / 2 (patch test); scaling all coordinates by h scales K
    by h."""
    from sprgpu import fem

    cube = fem.hex8_coordinates(1)
    X = cube[[0, 1, 3, 2, 4, 5, 7, 6]]  # the grid's nodes in local node order
    K0 = fem.hex8_stiffness(2.5, 0.3)
    K1 = fem.hex8_stiffness_at(X, 2.5, 0.3)
    assert np.max(np.abs(K1 - K0)) <= 1e-14 * np.max(np.abs(K0))
    rng = np.random.default_rng(4)
    Xd = X + rng.uniform(-0.2, 0.2, X.shape)
    K = fem.hex8_stiffness_at(Xd, 1.7, 0.25)
    assert np.max(np.abs(K - K.T)) <= 1e-14 * np.max(np.abs(K))
    w = np.linalg.eigvalsh(0.5 * (K + K.T))
    assert np.all(w[:6] < 1e-12 * w[-1]) and w[6] > 1e-6 * w[-1]
    rigid = [np.tile(e, 8) for e in np.eye(3)]
    for axis in np.eye(3):
        rigid.append(np.cross(axis, Xd).ravel())
    for r in rigid:
        assert np.max(np.abs(K @ r)) <= 1e-13 * np.max(np.abs(K)) * np.max(np.abs(r))
    # patch test on a parallelepiped (affine map of the cube): exact energy
    A = np.array([[1.2, 0.1, 0.0], [0.2, 0.9, 0.1], [0.0, 0.3, 1.1]])
    Xp = X @ A.T
    E, nu = 3.0, 0.3
    Kp = fem.hex8_stiffness_at(Xp, E, nu)
    G = rng.normal(size=(3, 3))
    u = (Xp @ G.T).ravel()
    eps = 0.5 * (G + G.T)
    lam, mu = E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
    vol = abs(np.linalg.det(A))
    energy = 0.5 * vol * (lam * np.trace(eps) ** 2 + 2 * mu * np.sum(eps * eps))
    assert abs(0.5 * u @ Kp @ u - energy) <= 1e-12 * energy
    assert np.max(np.abs(fem.hex8_stiffness_at(2 * Xd, 1.7, 0.25) - 2 * K)) <= 1e-13 * np.max(np.abs(K))
