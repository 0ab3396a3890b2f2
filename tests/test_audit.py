"""FP64 audit of a world (bp_world_audit, bp_batch_audit; the oracle's bpo_audit).

CPU: the oracle's brute-force audit against closed forms of the reference's Double oracle
(ContactLawReference.swift:26-31 energy, :68-82 pair of different materials, restated in
tests/contact_law_reference.py), against filters, and against the conservation the reference's
pair-bounce test states (NativeMetalSubstepKernelTests.swift:957-978: momentum to float accuracy,
energy within 2 % once the pair has separated). GPU: the device audit, which walks the rows of
the last grid build, against the oracle's brute force over all pairs, segments and bodies."""
import numpy as np
import pytest

import contact_law_reference as ref
import kernel_scenes as ks
from bubblephysics_amd import abi, scenes, solver


def two_bubbles(distance, ra=12.0, rb=12.0, ka=229.0, kb=229.0, va=(0.0, 0.0), vb=(0.0, 0.0), mask_b=0xFFFFFFFF):
    return ks.world([ks.body(1, (0.0, 0.0), velocity=va, radius=ra, stiffness=ka),
                     ks.body(2, (distance, 0.0), velocity=vb, radius=rb, stiffness=kb, mask=mask_b)])


def test_same_material_pair_energy_is_the_closed_form(oracle):
    for d in (23.0, 20.0, 12.5, 3.0):
        a = oracle.audit(two_bubbles(d))
        eps = (24.0 - np.float64(np.float32(d))) / 24.0
        assert a["contacts"] == 1
        assert np.isclose(a["contact_energy"], ref.energy(eps, 229.0, 24.0), rtol=1e-12)
    a = oracle.audit(two_bubbles(24.5))
    assert a["contacts"] == 0 and a["contact_energy"] == 0.0


def test_different_material_pair_energy_splits_the_overlap(oracle):
    for d, ka, kb, ra, rb in ((20.0, 229.0, 900.0, 12.0, 12.0), (15.0, 50.0, 229.0, 8.0, 14.0), (30.0, 1000.0, 10.0, 20.0, 16.0)):
        a = oracle.audit(two_bubbles(d, ra=ra, rb=rb, ka=ka, kb=kb))
        pc = ref.pair_contact(ra + rb - d, ra, ka, rb, kb)
        expected = ref.energy(pc.epsilon_a, ka, ra) + ref.energy(pc.epsilon_b, kb, rb)
        assert np.isclose(a["contact_energy"], expected, rtol=1e-9), (d, a["contact_energy"], expected)
        # A planted mistake: the whole overlap on the softer bubble is a different number.
        soft = ref.energy((ra + rb - d) / (ra if ka < kb else rb), min(ka, kb), ra if ka < kb else rb)
        assert not np.isclose(a["contact_energy"], soft, rtol=1e-3)


def test_wall_contact_energy_and_filters(oracle):
    w = ks.world([ks.body(1, (100.0, -5.0))], walls=True)           # 5 outside the one-sided bottom edge, radius 12
    a = oracle.audit(w)
    assert a["contacts"] == 1
    assert np.isclose(a["contact_energy"], ref.energy((12.0 - 5.0) / 12.0, 229.0, 12.0), rtol=1e-12)
    # A pair whose filters reject each other has no contact and no energy (Geometry:130-132).
    f = oracle.audit(two_bubbles(20.0, mask_b=0))
    assert f["contacts"] == 0 and f["contact_energy"] == 0.0


def test_momentum_and_kinetic_energy_closed_forms(oracle):
    w = two_bubbles(100.0, va=(600.0, -3.0), vb=(-400.0, 7.0))
    a = oracle.audit(w)
    m = np.float64(np.float32(0.36))
    assert np.allclose(a["momentum"], [m * 600 - m * 400, m * -3 + m * 7], rtol=1e-12)
    assert np.isclose(a["kinetic_energy"], 0.5 * m * (600**2 + 3**2 + 400**2 + 7**2), rtol=1e-12)
    assert np.isclose(a["angular_momentum"], m * (100.0 * 7.0), rtol=1e-12)


def test_pair_bounce_conserves_momentum_and_energy(oracle):
    # NativeMetalSubstepKernelTests.swift:957-978: masses 0.36 and 1.5, speeds 600 and -400.
    w = ks.world([ks.body(1, (0.0, 0.0), velocity=(600.0, 0.0), mass=0.36),
                  ks.body(2, (60.0, 0.0), velocity=(-400.0, 0.0), mass=1.5)])
    a0 = oracle.audit(w)
    e0 = a0["kinetic_energy"] + a0["contact_energy"]
    assert np.isclose(e0, 0.5 * 0.36 * 600**2 + 0.5 * 1.5 * 400**2, rtol=1e-6)
    cur = w.copy()
    touched = False
    for _ in range(12):
        cur.bubbles, _, frames, _ = oracle.run(cur, 1)
        a = oracle.audit(cur)
        touched |= a["contacts"] > 0
        assert abs(a["momentum"][0] - a0["momentum"][0]) <= 2e-3 * max(1.0, abs(a0["momentum"][0]))
        # Total energy, spring included, stays within the leapfrog error at every frame.
        assert abs(a["kinetic_energy"] + a["contact_energy"] - e0) / e0 <= 0.02
    assert touched and a["contacts"] == 0


@pytest.mark.gpu
def test_device_audit_matches_the_oracle_brute_force(oracle):
    scene = scenes.config1()
    world = solver.NativeWorld(scene)
    for k in range(10):
        world.step(k)
    got = world.audit()
    cb, cbod, _, _ = oracle.run(scene, 10)
    want = oracle.audit(scene, cb, cbod)
    assert got["bubbles"] == want["bubbles"] and got["contacts"] == want["contacts"] > 0
    for f in ("kinetic_energy", "contact_energy", "angular_momentum"):
        assert np.isclose(got[f], want[f], rtol=1e-10, atol=1e-9), f
    assert np.allclose(got["momentum"], want["momentum"], rtol=1e-10, atol=1e-9)


@pytest.mark.gpu
def test_batch_audit_with_bodies_matches_the_oracle(oracle):
    rng = np.random.Generator(np.random.PCG64(41))
    ws = [scenes.batch_world(rng, n, bodies_max=12) for n in (40, 700, 3000)]
    for w in ws:
        w.config.export_adjacency = 1
    quiet = ws[0].copy()
    quiet.config.export_adjacency = 0
    batch = solver.Batch(ws + [quiet])
    batch.step(4)
    got = batch.audit()
    for k, w in enumerate(ws):
        cb, cbod, frames, _ = oracle.run(w, 4)
        if frames[-1]["status"] != abi.BP_OK:
            continue
        want = oracle.audit(w, cb, cbod)
        assert got[k]["contacts"] == want["contacts"], k
        assert np.isclose(got[k]["contact_energy"], want["contact_energy"], rtol=1e-10, atol=1e-9), k
        assert np.isclose(got[k]["kinetic_energy"], want["kinetic_energy"], rtol=1e-10), k
    assert np.isnan(got[-1]["contact_energy"]) and got[-1]["contacts"] == 0
    assert np.isclose(got[-1]["kinetic_energy"], got[0]["kinetic_energy"], rtol=1e-12)
