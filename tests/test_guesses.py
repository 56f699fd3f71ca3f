"""Analytic initial configurations (glh_initial_guess_field): B-in-A ball,
B-n-B polka dots and the weak-coupling PDW ansatz, pinned by their defining
regions, special points where the lateral factors reduce to constants, the
lattice periods of the PDW pattern, linearity in the amplitude u, and the
noise stream shared with the seeded bulk guess."""
import numpy as np
import pytest

from product_util import glhip

G = glhip()

F32_B = float(np.float32(0.5773502691896258))
F32_A = float(np.float32(0.7071067811865475))


def grid(n=7, L=10.0, hz=6.0):
    x = np.linspace(-L, L, n)
    z = np.linspace(-hz, hz, n)
    return np.array([(a, b, c) for c in z for b in x for a in x])


def test_bina_ball_regions():
    xyz = grid()
    r, gA, gB = 6.0, 1.3, 1.7
    U = G.initial_guess_field("BinA", xyz, (r, gA, gB)).reshape(-1, 18)
    inside = (xyz ** 2).sum(1) <= r * r
    assert inside.any() and (~inside).any()
    B = np.zeros(18); B[[0, 4, 8]] = gB * F32_B
    A = np.zeros(18); A[[0, 10]] = gA * F32_A
    assert np.array_equal(U[inside], np.tile(B, (inside.sum(), 1)))
    assert np.array_equal(U[~inside], np.tile(A, ((~inside).sum(), 1)))


def test_bnb_cylinder_regions():
    xyz = grid()
    U = G.initial_guess_field("BnB", xyz, (5.0,)).reshape(-1, 18)
    inside = xyz[:, 0] ** 2 + xyz[:, 1] ** 2 <= 25.0
    assert np.array_equal(U[:, :9], U[:, 9:])        # u = v
    assert U[inside, 0][0] == 0.859734 and U[~inside, 0][0] == 0.668541
    assert U[inside, 8][0] == 0.245094 and U[~inside, 8][0] == -0.799212
    # constant over z within each region
    assert np.all(U[inside] == U[inside][0]) and np.all(U[~inside] == U[~inside][0])


def test_pdw_special_points():
    # at x = y = 0: cos terms 1, sin terms 0, so cos(qx) - cos(qx/2)cos(..) = 0
    # and cos(qx) + 2cos(qx/2)cos(..) = 3; the z profiles at s = 0 are their
    # constant coefficients, at s = 1 their coefficient sums
    u, hD, gap = 0.7, 6.0, 1.9
    p6 = np.array([-0.010674533968153196, -0.04928146307995591, 0.4123298769345518, -5.957210687551436,
                   15.697110273856326, -14.963347649994844, 4.860888318571941])
    for z, P6 in ((-hD, p6[0]), (hD, p6.sum())):
        U = G.initial_guess_field("PDW", np.array([[0.0, 0.0, z]]), (u, hD, gap))
        g, b0 = gap * u, gap / np.sqrt(3.0)
        np.testing.assert_allclose(U[0], b0 + 2 * g * P6, rtol=1e-13, atol=1e-14)
        np.testing.assert_allclose(U[1], 2 * g * P6, rtol=1e-13, atol=1e-14)
        np.testing.assert_allclose(U[2], 2 * g * P6, rtol=1e-13, atol=1e-14)
        np.testing.assert_allclose(U[3], 6 * np.sqrt(3) * g * P6, rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(U[4], b0 + 6 * np.sqrt(3) * g * P6, rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(U[5], 6 * np.sqrt(3) * g * P6, rtol=1e-12, atol=1e-14)
        assert abs(U[6]) < 1e-15 and abs(U[7]) < 1e-15 and U[8] == b0
        assert np.all(U[9:] == 0.0)


def test_pdw_lattice_periods_and_linearity():
    hD = 6.0
    q = np.pi / (np.sqrt(3.0) * 2.0 * hD)
    rng = np.random.default_rng(0)
    xyz = np.column_stack([rng.uniform(-20, 20, 50), rng.uniform(-20, 20, 50), rng.uniform(-hD, hD, 50)])
    U = G.initial_guess_field("PDW", xyz, (1.0, hD, 1.0))
    for shift in ((4 * np.pi / q, 0, 0), (0, 4 * np.pi / (np.sqrt(3) * q), 0)):
        Us = G.initial_guess_field("PDW", xyz + np.array(shift), (1.0, hD, 1.0))
        np.testing.assert_allclose(Us, U, rtol=0, atol=1e-9)
    # a half period in x flips the sign of the sin(qx/2) terms only: not a symmetry of the pattern
    Uh = G.initial_guess_field("PDW", xyz + np.array([2 * np.pi / q, 0, 0]), (1.0, hD, 1.0))
    assert np.abs(Uh - U).max() > 1e-3
    b0 = np.zeros(18); b0[[0, 4, 8]] = 1.0 / np.sqrt(3.0)
    U2 = G.initial_guess_field("PDW", xyz, (2.5, hD, 1.0))
    np.testing.assert_allclose((U2.reshape(-1, 18) - b0), 2.5 * (U.reshape(-1, 18) - b0), rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("kind,par,scale", [("BinA", (4.0, 1.0, 2.0), 2.0), ("BnB", (3.0,), 1.0),
                                            ("PDW", (1.0, 6.0, 3.0), 3.0)])
def test_noise_is_the_seeded_stream(kind, par, scale):
    xyz = grid(5)
    n = xyz.shape[0]
    U0 = G.initial_guess_field(kind, xyz, par)
    U1 = G.initial_guess_field(kind, xyz, par, amp=0.05, seed=7)
    N = G.initial_guess(n, scale, phase=0, amp=0.05, seed=7) - G.initial_guess(n, scale, phase=0, amp=0.0, seed=7)
    np.testing.assert_allclose(U1 - U0, N, rtol=0, atol=1e-14)
