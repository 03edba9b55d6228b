"""The CPU oracle against what the mathematics fixes (closed forms, invariants,
brute force), independently of the reference fixtures.  Each check is chosen
so that a plausible slip (a dropped term, a sign, a transposed operand, a wrong
neighbour) fails it."""
import numpy as np
import pytest

import gfc
import oracle as orc

RNG = np.random.default_rng(2024)


def rand_su3(n=1):
    out = []
    for _ in range(n):
        z = RNG.normal(size=(3, 3)) + 1j * RNG.normal(size=(3, 3))
        q, r = np.linalg.qr(z)
        q = q @ np.diag(np.diag(r) / np.abs(np.diag(r)))
        q /= np.linalg.det(q) ** (1 / 3)
        out.append(q)
    return out if n > 1 else out[0]


def is_su3(m, tol=1e-13):
    return (np.max(np.abs(m @ m.conj().T - np.eye(3))) < tol
            and abs(np.linalg.det(m) - 1) < tol)


def test_product_and_inverse_against_numpy():
    for _ in range(50):
        a = RNG.normal(size=(3, 3)) + 1j * RNG.normal(size=(3, 3))
        b = RNG.normal(size=(3, 3)) + 1j * RNG.normal(size=(3, 3))
        assert np.allclose(orc.prod(a, b), a @ b, atol=1e-14)
        assert np.allclose(orc.inverse(a) @ a, np.eye(3), atol=1e-12)
        assert abs(orc.det(a) - np.linalg.det(a)) < 1e-13 * max(1, abs(np.linalg.det(a)))


def test_project_su3_properties():
    for _ in range(50):
        a = RNG.normal(size=(3, 3)) + 1j * RNG.normal(size=(3, 3))
        st, p = orc.project_su3(a)
        assert st == 0 and is_su3(p)
        # row 0 keeps its direction
        r0 = a[0] / np.linalg.norm(a[0])
        assert np.allclose(p[0], r0, atol=1e-14)
        # idempotent on SU(3)
        assert np.allclose(orc.project_su3(p)[1], p, atol=1e-14)
    assert orc.project_su3(np.zeros((3, 3)))[0] == -1


def test_cm_step_maximises_subgroup_trace():
    """Each Cabibbo-Marinari step left-multiplies rows (a, b) by the SU(2)
    element maximising Re Tr(x w_sub); brute force over random SU(2)."""
    pairs = [(0, 1), (0, 2), (1, 2)]
    for _ in range(20):
        w = RNG.normal(size=(3, 3)) + 1j * RNG.normal(size=(3, 3))
        for sub, (a, b) in enumerate(pairs):
            w2 = orc.cm_step(w, sub)
            x = w2[[a, b]] @ np.linalg.pinv(w[[a, b]])
            # x is in SU(2)
            assert np.allclose(x @ x.conj().T, np.eye(2), atol=1e-12)
            assert abs(np.linalg.det(x) - 1) < 1e-12
            best = np.trace(x @ w[np.ix_([a, b], [a, b])]).real
            for _ in range(200):
                v = RNG.normal(size=4)
                v /= np.linalg.norm(v)
                y = np.array([[v[0] + 1j * v[3], v[2] + 1j * v[1]],
                              [-v[2] + 1j * v[1], v[0] - 1j * v[3]]])
                assert np.trace(y @ w[np.ix_([a, b], [a, b])]).real <= best + 1e-12
            # rows outside (a, b) untouched
            c = 3 - a - b
            assert np.array_equal(w2[c], w[c])


def test_la_update_equals_product_of_su2_factors():
    w = sum(rand_su3(6))
    g = orc.la_update(w, 2)
    assert is_su3(g, 1e-12)
    # w_final = g w and it has a larger real trace than w
    assert np.trace(g @ w).real > np.trace(w).real
    # degenerate branch: det(w) = 0 -> g = 0
    assert np.array_equal(orc.la_update(np.zeros((3, 3)), 2), np.zeros((3, 3)))


def test_power_binomial_closed_forms():
    g = rand_su3()
    st, p = orc.power_binomial(g, 1.0)
    assert st == 0 and np.allclose(p, orc.project_su3(g)[1], atol=1e-14)
    # near the identity: Proj(1 + w(A - 1)) ~ exp(i w eps X)
    h = RNG.normal(size=(3, 3)) + 1j * RNG.normal(size=(3, 3))
    h = (h + h.conj().T) / 2
    h -= np.trace(h) / 3 * np.eye(3)
    eps, om = 1e-4, 1.9
    w_, v = np.linalg.eigh(h)
    A = v @ np.diag(np.exp(1j * eps * w_)) @ v.conj().T
    B = v @ np.diag(np.exp(1j * om * eps * w_)) @ v.conj().T
    st, p = orc.power_binomial(A, om)
    assert np.max(np.abs(p - B)) < 50 * eps ** 2
    # degenerate: zero update -> Proj((1 - w) 1) = diag(-1, -1, 1)
    st, p = orc.power_binomial(np.zeros((3, 3)), 1.95)
    assert np.allclose(p, np.diag([-1, -1, 1]))


def test_decompose_gell_mann():
    lam = np.zeros((9, 3, 3), complex)
    lam[1][0, 1] = lam[1][1, 0] = 1
    lam[2][0, 1], lam[2][1, 0] = -1j, 1j
    lam[3][0, 0], lam[3][1, 1] = 1, -1
    lam[4][0, 2] = lam[4][2, 0] = 1
    lam[5][0, 2], lam[5][2, 0] = -1j, 1j
    lam[6][1, 2] = lam[6][2, 1] = 1
    lam[7][1, 2], lam[7][2, 1] = -1j, 1j
    lam[8] = np.diag([1, 1, -2]) / np.sqrt(3)
    c = RNG.normal(size=9)
    c[0] = 0
    A = sum(c[a] * lam[a] / 2 for a in range(1, 9))
    got = orc.decompose(A)
    assert np.allclose(got[1:], c[1:], atol=1e-14)  # c_a = Tr(lambda_a A)


def _lattice(ns, nt, eps, seed):
    geom = gfc.Geometry(ns, nt)
    U = gfc.generate_smooth(geom, seed, eps)
    return geom, U, orc.geom(ns, nt)


def test_residues_of_trivial_and_pure_gauge_fields():
    geom, U, g = _lattice(4, 4, 0.0, 3)
    one = np.zeros_like(U)
    one[:] = np.eye(3)
    assert orc.e2(one, g) == 0 and orc.theta(one, g) == 0
    assert abs(orc.functional(one, g) - 1) < 1e-15
    assert abs(orc.plaquette(one, g, 0) - 1) < 1e-15
    # theta = e2 / 6 identically (sum of |A_ab|^2 = (1/2) sum c_a^2, /3 vs /1)
    geom, U, g = _lattice(4, 4, 0.3, 4)
    assert abs(orc.theta(U, g) - orc.e2(U, g) / 6) < 1e-13 * orc.e2(U, g)


def test_plaquette_gauge_invariance_and_transform():
    geom, U, g = _lattice(4, 4, 0.3, 5)
    G = np.array(rand_su3(geom.volume))
    V = U.copy()
    orc.apply_gauge(V, G, g)
    for w in (0, 1):
        assert abs(orc.plaquette(V, g, w) - orc.plaquette(U, g, w)) < 1e-14
    # the transform is the definition U_mu(n) -> G(n) U G(n+mu)^dagger
    s, mu = 37, 2
    t, z, y, x = np.unravel_index(s, (4, 4, 4, 4))
    n2 = np.ravel_multi_index((t, z, (y + 1) % 4, x), (4, 4, 4, 4))
    assert np.allclose(V[s, mu], G[s] @ U[s, mu] @ G[n2].conj().T, atol=1e-14)


def test_sweep_invariants():
    """A half-sweep keeps plaquettes, keeps U = G U0 G(n+mu)^dagger, keeps G in
    SU(3); with omega = 1 (pure Los Alamos) F cannot decrease."""
    geom, U0, g = _lattice(4, 8, 0.25, 6)
    orc.reunitarize(U0)
    for omega in (1.0, 1.95):
        U = U0.copy()
        G = np.zeros((geom.volume, 3, 3), complex)
        G[:] = np.eye(3)
        F = [orc.functional(U, g)]
        for s in range(6):
            orc.half_sweep(U, G, g, s, omega, 2)
            F.append(orc.functional(U, g))
        V = U0.copy()
        orc.apply_gauge(V, G, g)
        assert np.max(np.abs(V - U)) < 1e-13
        assert all(is_su3(m, 1e-13) for m in G[:50])
        for w in (0, 1):
            assert abs(orc.plaquette(U, g, w) - orc.plaquette(U0, g, w)) < 1e-14
        if omega == 1.0:
            assert all(b >= a - 1e-14 for a, b in zip(F, F[1:]))


def test_pure_gauge_converges_to_trivial_gauge():
    """eps = 0: U is a pure gauge, so e2 -> 0 and F -> 1 (no Gribov copy on
    this small smooth case)."""
    geom, U, g = _lattice(4, 4, 0.0, 7)
    orc.reunitarize(U)
    G = np.zeros((geom.volume, 3, 3), complex)
    rc, _ = orc.gaugefix(U, G, g, orc.params(tolerance=1e-20, estimate=0, reunit_every=10))
    assert rc > 0 and orc.e2(U, g) <= 1e-20
    # spatial links become constant within each timeslice (global residual
    # gauge freedom per slice), so the spatial plaquette is 1 and F -> 1
    assert abs(orc.functional(U, g) - 1) < 1e-12


def test_check_schedule_and_stall():
    p = orc.params(tolerance=1e-16, estimate=1000)
    assert orc.next_check(0, 16.0, p) == 10 + int(1000 * (1 + np.log10(16) / 16))
    assert orc.next_check(5, 1e-17, p) == 5
    # hits = 0, omega = 1: nothing changes, the first check stalls -> -1
    geom, U, g = _lattice(4, 4, 0.3, 8)
    G = np.zeros((geom.volume, 3, 3), complex)
    rc, _ = orc.gaugefix(U, G, g, orc.params(estimate=0, hits=0, omega=1.0))
    assert rc == -1
    assert orc.gaugefix(U, G, g, orc.params(omega=2.0))[0] == -2
