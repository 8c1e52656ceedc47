"""Pins for the FP64 CPU oracle (oracle/imex_oracle.c).

Each check is chosen so that a plausible slip in the restatement (a dropped
equilibrium term, a wrong sign, a transposed index, a wrong relaxation rate)
fails it, and none of them retypes the oracle's formulas: they use closed-form
lattice moments, exact permutations (numpy.roll), the reference's own quoted
numbers (tests/regression.sh:28-43) and hand-evaluated decompositions
(src/mpi_view_transfer.cpp:55-91).
"""
import math

import numpy as np
import pytest

from tests.conftest import REGRESSION, ke_matches_reference

PI_LIT = 3.1415926  # src/lbm.cpp:3


def test_lattice_moments(oracle_mod):
    w, e, opp = oracle_mod.tables()
    assert w.sum() == pytest.approx(1.0, abs=4e-16)
    assert np.abs(w @ e).max() < 1e-16
    m2 = np.einsum("i,ia,ib->ab", w, e, e)
    assert np.allclose(m2, np.eye(3) / 3.0, atol=2e-16, rtol=0)
    m4 = np.einsum("i,ia,ib,ic,id->abcd", w, e, e, e, e)
    d = np.eye(3)
    iso = (np.einsum("ab,cd->abcd", d, d) + np.einsum("ac,bd->abcd", d, d)
           + np.einsum("ad,bc->abcd", d, d)) / 9.0
    assert np.allclose(m4, iso, atol=2e-16, rtol=0)
    assert np.array_equal(e[opp], -e)
    # order: rest, 6 faces, 12 edges, 8 corners (src/lbm.cpp:20-30)
    nnz = (e != 0).sum(axis=1)
    assert list(nnz) == [0] + [1] * 6 + [2] * 12 + [3] * 8
    assert list(e[1]) == [1, 0, 0] and list(e[19]) == [1, 1, 1]


def test_tau_closed_form(oracle_mod):
    # tau0 = u0*sx/Re/cs2/2/3.1415926 (src/System.cpp:28)
    assert oracle_mod.tau(0.05, 32, 1600) == pytest.approx(0.0015 / PI_LIT, rel=2e-16)
    om = lambda t: 1.0 / (t + 0.5)
    assert om(oracle_mod.tau(0.05, 32, 100)) == pytest.approx(1.970, abs=5e-4)
    assert om(oracle_mod.tau(0.05, 256, 1600)) == pytest.approx(1.985, abs=5e-4)
    assert om(oracle_mod.tau(0.08, 640, 1600)) == pytest.approx(1.941, abs=5e-4)


@pytest.mark.parametrize("n,expect", [(1, (1, 1, 1)), (2, (1, 1, 2)), (4, (1, 2, 2)),
                                      (8, (2, 2, 2))])
def test_decompose_hand_evaluated(oracle_mod, n, expect):
    assert oracle_mod.decompose(n, 32, 32, 32) == expect


def test_block_geometry(oracle_mod):
    o = oracle_mod.Oracle(17, 17, 17, 0.01, 1.0, 0.05, nranks=4)
    assert o.grid() == (1, 2, 2)
    # ranks with coordinate < gl % r own one extra cell (src/mpi_view_transfer.cpp:185-188)
    geo = [o.rank_geometry(m) for m in range(4)]
    # me = px + rx*pz + rx*rz*py: rank 1 is (0,0,1), rank 2 is (0,1,0)
    assert geo[1]["p"] == (0, 0, 1) and geo[2]["p"] == (0, 1, 0)
    assert geo[0]["l"] == (17, 9, 9) and geo[3]["l"] == (17, 8, 8)
    assert geo[3]["lo"] == (0, 9, 9)
    o.close()


def _analytic(n, u0, rho0):
    x = np.arange(n) / n * 2.0 * PI_LIT
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    u = u0 * np.sin(X) * np.cos(Y) * np.cos(Z)
    v = -u0 * np.cos(X) * np.sin(Y) * np.cos(Z)
    p = rho0 / 3.0 + rho0 * u0 * u0 / 16.0 * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2.0)
    return u, v, p


def test_initial_state_moments(oracle_mod):
    n, u0, rho0 = 16, 0.05, 1.0
    o = oracle_mod.Oracle(n, n, n, 0.01, rho0, u0)
    o.initialize()
    f = o.get_f()
    w, e, _ = oracle_mod.tables()
    u, v, p = _analytic(n, u0, rho0)
    rho = f.sum(axis=0)
    mom = np.einsum("ia,ixyz->axyz", e.astype(float), f)
    assert np.abs(rho - 3 * p).max() < 1e-15
    assert np.abs(mom[0] - u).max() < 1e-16 and np.abs(mom[1] - v).max() < 1e-16
    assert np.abs(mom[2]).max() < 1e-16
    # second moment of the equilibrium: p delta_ab + u_a u_b
    pi = np.einsum("ia,ib,ixyz->abxyz", e.astype(float), e.astype(float), f)
    uu = np.stack([u, v, np.zeros_like(u)])
    want = np.einsum("ab,xyz->abxyz", np.eye(3), p) + np.einsum("axyz,bxyz->abxyz", uu, uu)
    assert np.abs(pi - want).max() < 1e-15
    c = o.conserved()
    assert c[0] == pytest.approx(3 * p.sum(), rel=1e-14)
    # KE_0 = sum 0.5|u|^2 -> N u0^2 / 8 on a periodic grid
    assert c[4] == pytest.approx(n ** 3 * u0 ** 2 / 8.0, rel=2e-6)
    o.close()


def test_collision_relaxes_nonequilibrium_moments(oracle_mod):
    n, rho0 = 4, 1.0
    tau0 = 0.037
    omega = 1.0 / (tau0 + 0.5)
    w, e, _ = oracle_mod.tables()
    rng = np.random.default_rng(7)
    f0 = w[:, None, None, None] * rho0 * (1.0 + 0.2 * rng.standard_normal((27, n, n, n)))
    o = oracle_mod.Oracle(n, n, n, tau0, rho0, 0.05)
    o.set_f(f0)
    o.collision()
    f1 = o.get_f()
    ef = e.astype(float)
    rho0_, rho1 = f0.sum(0), f1.sum(0)
    j0, j1 = np.einsum("ia,ixyz->axyz", ef, f0), np.einsum("ia,ixyz->axyz", ef, f1)
    assert np.abs(rho1 - rho0_).max() < 1e-15
    assert np.abs(j1 - j0).max() < 1e-15
    # Pi_eq = p delta + u u with p = rho/3, u = j (pressure-based equilibrium)
    pi0 = np.einsum("ia,ib,ixyz->abxyz", ef, ef, f0)
    pi1 = np.einsum("ia,ib,ixyz->abxyz", ef, ef, f1)
    peq = np.einsum("ab,xyz->abxyz", np.eye(3), rho0_ / 3.0) + np.einsum("axyz,bxyz->abxyz", j0, j0)
    assert np.abs((pi1 - peq) - (1.0 - omega) * (pi0 - peq)).max() < 2e-15
    o.close()


def test_collision_equilibrium_fixed_point(oracle_mod):
    n = 8
    o = oracle_mod.Oracle(n, n, n, oracle_mod.tau(0.05, n, 100), 1.0, 0.05)
    o.initialize()
    f0 = o.get_f()
    o.collision()
    assert np.abs(o.get_f() - f0).max() < 2e-15
    o.close()


@pytest.mark.parametrize("nranks", [1, 2, 8])
def test_streaming_is_the_periodic_pull_permutation(oracle_mod, nranks):
    nx, ny, nz = 6, 5, 7
    w, e, _ = oracle_mod.tables()
    tag = np.arange(27 * nx * ny * nz, dtype=np.float64).reshape(27, nx, ny, nz) + 0.5
    o = oracle_mod.Oracle(nx, ny, nz, 0.01, 1.0, 0.05, nranks=nranks)
    o.set_f(tag)
    o.streaming()
    got = o.get_f()
    for ii in range(27):
        want = np.roll(tag[ii], shift=tuple(e[ii]), axis=(0, 1, 2))
        assert np.array_equal(got[ii], want), ii
    o.close()


@pytest.mark.parametrize("nranks", [2, 4, 8])
def test_decomposition_independence_bitwise(oracle_mod, nranks):
    # tests/regression.sh:17-24: populations do not depend on the split.
    n = 17
    t = oracle_mod.tau(0.05, n, 400)
    ref = oracle_mod.Oracle(n, n, n, t, 1.0, 0.05, 1)
    ref.initialize(); ref.step(6)
    o = oracle_mod.Oracle(n, n, n, t, 1.0, 0.05, nranks)
    o.initialize(); o.step(6)
    assert np.array_equal(ref.get_f(), o.get_f())
    a, b = ref.conserved(), o.conserved()
    assert abs(a[0] - b[0]) / a[0] < 1e-12 and abs(a[4] - b[4]) / a[4] < 1e-12
    ref.close(); o.close()


def test_regression_deck_known_answers(oracle_mod):
    # Deck of tests/regression.sh:60: 1.0 0.1 1600 / 0.05 20 1 / 32 32 32.
    n, re, u0 = 32, 1600, 0.05
    o = oracle_mod.Oracle(n, n, n, oracle_mod.tau(u0, n, re), 1.0, u0)
    o.initialize()
    cons = []
    for _ in range(20):
        o.step(1)
        cons.append(o.conserved())
    cons = np.array(cons)
    # KE transient quoted at tests/regression.sh:28-30
    assert ke_matches_reference(cons[:, 4])
    assert round(cons[1, 4], 3) == REGRESSION["ke_first_reports"][1]
    m0 = cons[0, 0]
    assert np.abs((cons[:, 0] - m0) / m0).max() < REGRESSION["tol_mass_drift_rel"]
    assert np.abs(cons[:, 1:4]).max() < REGRESSION["tol_momentum_abs"]
    o.close()


def test_conserved_matches_compensated_sum(oracle_mod):
    n = 24
    o = oracle_mod.Oracle(n, n, n, oracle_mod.tau(0.05, n, 100), 1.0, 0.05)
    o.initialize(); o.step(5)
    a, b = o.conserved(), o.conserved_compensated()
    assert abs(a[0] - b[0]) / b[0] < 1e-13
    assert abs(a[4] - b[4]) / b[4] < 1e-12
    assert np.abs(a[1:4] - b[1:4]).max() < 1e-12
    o.close()


def test_viscous_decay_rate(oracle_mod):
    # nu = cs2 * tau0 for the rate 1/(tau0+0.5); at low Re the TGV kinetic
    # energy decays as exp(-2 nu k^2 t) with k^2 = 3 / L^2, L = n / (2 pi').
    # The lattice rate approaches it as the grid is refined.
    def ratio(n, re=5, u0=0.02):
        t = oracle_mod.tau(u0, n, re)
        nu = t / 3.0
        L = n / (2.0 * PI_LIT)
        want = 6.0 * nu / L ** 2
        half = int(0.3 / want)
        o = oracle_mod.Oracle(n, n, n, t, 1.0, u0)
        o.initialize()
        o.step(half)
        k1 = o.conserved()[4]
        o.step(half)
        k2 = o.conserved()[4]
        o.close()
        return math.log(k1 / k2) / half / want

    r16, r32 = ratio(16), ratio(32)
    assert abs(r32 - 1.0) < 0.02
    assert abs(r32 - 1.0) < 0.6 * abs(r16 - 1.0)
