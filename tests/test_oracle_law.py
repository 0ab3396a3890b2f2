"""Pins the oracle's contact law (NativeForceKernels.metal-source:28-102) against the Double
reference of the reference's tests and against closed forms, with the tolerances of
Tests/BubblePhysicsNativeMetalTests/NativeMetalContactLawTests.swift."""
import math

import numpy as np
import pytest

import contact_law_reference as ref


def rel(value, expected):
    return abs(value - expected) / max(abs(expected), 2.2250738585072014e-308)


def test_inverse_law_float32_matches_double_from_1e_minus_4_to_1e6(oracle):
    # NativeMetalContactLawTests.swift:74-90: 601 log-spaced q, relative error <= 1e-3, eps < 1.
    q = np.array([10.0 ** (-4.0 + 10.0 * k / 600) for k in range(601)], np.float32)
    eps = oracle.epsilon(q)
    errors = [rel(float(e), ref.epsilon_for_ratio(float(x))) for e, x in zip(eps, q)]
    assert max(errors) <= 1e-3
    assert np.all(eps < 1)


def test_inverse_law_is_the_inverse_of_the_force_law(oracle):
    # :92-102: F(eps)/K / q == 1 within 2e-3, eps rounded to float32.
    q = np.array([1e-4, 1e-3, 0.05, 1, 40, 1e3, 1e6], np.float32)
    for value, e in zip(q, oracle.epsilon(q)):
        assert ref.force(float(e), 1.0) / float(value) == pytest.approx(1.0, abs=2e-3)


def wall_cases():
    cases = []
    for stiffness, radius in [(229.0, 24.0), (22.0, 148.0), (22900.0, 12.0), (1.5, np.float32(7168) / np.float32(48))]:
        for fraction in [0.999, 0.99, 0.9, 0.75, 0.5, 0.25, 0.1, 0.03, 0.01, 0.004, 0.002]:
            r = np.float32(radius)
            cases.append((np.float32(stiffness), r, np.float32(r * np.float32(fraction))))
    return cases


def test_wall_law_matches_force_and_stiffness_in_double(oracle):
    # :106-132: F within 1e-4 relative, eps within 1e-5, k = dF/dx within 1e-3, second eps 0.
    cases = wall_cases()
    out = oracle.wall_law([(k, r, d, 0) for k, r, d in cases])
    for (k, r, d), (f, eps, stiff, eps_b) in zip(cases, out):
        epsilon = 1 - float(d) / float(r)
        expected = ref.force(epsilon, float(k))
        h = 1e-6 * (1 - epsilon)
        slope = (ref.force(epsilon + h, float(k)) - ref.force(epsilon - h, float(k))) / (2 * h * float(r))
        assert rel(float(f), expected) <= 1e-4
        assert abs(float(eps) - epsilon) <= 1e-5
        assert rel(float(stiff), slope) <= 1e-3
        assert eps_b == 0


def test_wall_law_zero_without_contact_and_finite_at_the_surface(oracle):
    # :134-147.
    out = oracle.wall_law([(229, 24, 24, 0), (229, 24, 30, 0), (229, 24, 0, 0), (229, 24, -3, 0), (0, 24, 10, 0)])
    assert out[0, 0] == 0 and out[1, 0] == 0 and out[4, 0] == 0
    for i in (2, 3):
        assert np.isfinite(out[i, 0]) and out[i, 0] > 1e6
        assert np.isfinite(out[i, 2])


def pair_cases():
    radii = [(24, 24), (24, 148), (148, 24), (12, 148), (60, 60)]
    stiff = [(229, 229), (229, 343.5), (229, 2290), (229, 229000), (229000, 229), (22, 22000)]
    fractions = [1e-3, 0.01, 0.05, 0.2, 0.5, 0.8, 0.95, 0.99, 0.998]
    return [(ra, rb, ka, kb, f) for ra, rb in radii for ka, kb in stiff for f in fractions]


def test_pair_law_matches_double_solution_same_and_different_stiffness(oracle):
    # :166-194: F within 1e-2 (worst 2e-3), deflection sum within 1e-3 R, eps within 1e-3.
    cases = pair_cases()
    inputs = [(ka, ra, kb, rb) for ra, rb, ka, kb, _ in cases]
    distances = [np.float32((ra + rb) * (1 - f)) for ra, rb, _, _, f in cases]
    out = oracle.pair_law(inputs, distances)
    worst = 0.0
    for (ra, rb, ka, kb, _), d, (force, ea, eb, _k) in zip(cases, distances, out):
        overlap = ra + rb - float(d)
        exact = ref.pair_contact(overlap, ra, ka, rb, kb)
        err = rel(float(force), exact.force)
        worst = max(worst, err)
        assert err <= 1e-2
        assert abs(ra * float(ea) + rb * float(eb) - overlap) <= 1e-3 * (ra + rb)
        assert abs(float(ea) - exact.epsilon_a) <= 1e-3
        assert abs(float(eb) - exact.epsilon_b) <= 1e-3
    assert worst <= 2e-3


def test_same_material_pair_uses_the_closed_form(oracle):
    # :196-218: F = K (R - d) R / d^2 within 1e-5, eps_a == eps_b, k = K (1+eps) / (R (1-eps)^3) within 2e-3.
    inputs, distances = [], []
    for ra, rb in [(24, 24), (24, 148), (148, 24), (12, 60)]:
        for f in [0.01, 0.2, 0.7]:
            inputs.append((229, ra, 229, rb))
            distances.append(np.float32((np.float32(ra) + np.float32(rb)) * (np.float32(1) - np.float32(f))))
    out = oracle.pair_law(inputs, distances)
    for (_, ra, _, rb), d, (force, ea, eb, k) in zip(inputs, distances, out):
        R, dd = ra + rb, float(d)
        assert rel(float(force), 229 * (R - dd) * R / (dd * dd)) <= 1e-5
        assert ea == eb
        e = float(ea)
        assert rel(float(k), 229 * (1 + e) / (R * (1 - e) ** 3)) <= 2e-3


def test_pair_stiffness_is_the_derivative_of_the_force(oracle):
    # :220-236.
    fractions = [0.05, 0.3, 0.7, 0.95]
    distances = [np.float32(172 * (1 - f)) for f in fractions]
    out = oracle.pair_law([(229, 24, 5000, 148)] * 4, distances)
    for d, row in zip(distances, out):
        overlap = 172 - float(d)
        h = 1e-5 * overlap
        up = ref.pair_contact(overlap + h, 24, 229, 148, 5000)
        down = ref.pair_contact(overlap - h, 24, 229, 148, 5000)
        assert rel(float(row[3]), (up.force - down.force) / (2 * h)) <= 1e-2


def test_pair_law_zero_without_overlap_and_finite_when_centres_coincide(oracle):
    # :238-251.
    inputs = [(229, 24, 2290, 24)] * 4 + [(229, 24, 229, 24)]
    out = oracle.pair_law(inputs, [48, 60, 0, 1e-9, 48])
    assert out[0, 0] == 0 and out[1, 0] == 0 and out[4, 0] == 0
    for i in (2, 3):
        assert np.isfinite(out[i, 0]) and out[i, 0] > 1e6
        assert np.all(np.isfinite(out[i, 1:]))


def test_planted_mistakes_are_caught_by_these_checks():
    """The closed forms above reject plausible slips of the law (sign, missing R, square)."""
    R, d, K = 48.0, 40.0, 229.0
    good = K * (R - d) * R / (d * d)
    for wrong in (K * (R - d) / (d * d), K * (d - R) * R / (d * d), K * (R - d) * R / d):
        assert rel(wrong, good) > 1e-3
    # epsilon with the cancelling form of the inverse loses digits at small q.
    q = 1e-3
    cancelling = np.float32(((2 * np.float32(q) + 1) - np.sqrt(np.float32(4 * q + 1))) / (2 * np.float32(q)))
    assert rel(float(cancelling), ref.epsilon_for_ratio(q)) > 1e-3


def test_half_impulse_factor_exact_and_series(oracle):
    # Force:754-758: (1 - e^{-x}) / gamma, series below x = 0.05; gamma = 0 gives h/2 exactly.
    gammas = np.array([0.0, 0.0, 1.0, 10.0, 100.0, 6.0], np.float32)
    halves = np.array([1 / 120, 1 / 240, 1 / 120, 1 / 120, 1 / 120, 0.01], np.float32)
    out = oracle.half_impulse_factor(gammas, halves)
    assert out[0] == halves[0] and out[1] == halves[1]
    for g, h, v in zip(gammas[2:], halves[2:], out[2:]):
        exact = -math.expm1(-float(g) * float(h)) / float(g)
        assert rel(float(v), exact) <= 2e-6


def test_deterministic_transcendentals_track_libm(oracle):
    x = np.linspace(-20.0, 20.0, 20001, dtype=np.float32)
    s, c = oracle.detmath("sin", x), oracle.detmath("cos", x)
    assert np.max(np.abs(s.astype(np.float64) - np.sin(x.astype(np.float64)))) <= 4e-7
    assert np.max(np.abs(c.astype(np.float64) - np.cos(x.astype(np.float64)))) <= 4e-7
    e_in = np.linspace(-80, 80, 10001, dtype=np.float32)
    e = oracle.detmath("exp", e_in).astype(np.float64)
    assert np.max(np.abs(e / np.exp(e_in.astype(np.float64)) - 1)) <= 3e-7
    l_in = np.exp(np.linspace(-60, 60, 10001)).astype(np.float32)
    lg = oracle.detmath("log", l_in).astype(np.float64)
    assert np.max(np.abs(lg - np.log(l_in.astype(np.float64)))) <= 2e-6
    assert oracle.detmath("exp", np.array([0.0, -0.0], np.float32)).tolist() == [1.0, 1.0]
    assert oracle.detmath("cos", np.array([0.0], np.float32))[0] == 1.0
    assert oracle.detmath("sin", np.array([0.0], np.float32))[0] == 0.0
