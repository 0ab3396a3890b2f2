"""Double-precision restatement of the reference's test oracle for the 4.0.0 contact law.

Follows Tests/BubblePhysicsRuntimeTests/ContactLawReference.swift:10-218 (the same functions,
written from its formulas): the law F = K eps / (1 - eps)^2, its inverse, the spring energy, the
bisection solution of a pair of different materials, and the 1D bubble-against-wall scene with the
h rule. It does not use the oracle or the HIP library, so it checks both independently.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field


def force(epsilon: float, stiffness: float) -> float:          # :14-17
    if not epsilon > 0:
        return 0.0
    return stiffness * epsilon / ((1 - epsilon) * (1 - epsilon))


def epsilon_for_ratio(q: float) -> float:                        # :21-24
    if not q > 0:
        return 0.0
    return 2 * q / ((2 * q + 1) + math.sqrt(4 * q + 1))


def energy(epsilon: float, stiffness: float, radius: float) -> float:   # :27-31
    if not epsilon > 0:
        return 0.0
    return stiffness * radius * (epsilon / (1 - epsilon) + math.log1p(-epsilon))


@dataclass
class PairContact:
    force: float
    epsilon_a: float
    epsilon_b: float


def pair_contact(overlap: float, ra: float, ka: float, rb: float, kb: float) -> PairContact | None:   # :68-82
    if not (overlap >= 0 and overlap < ra + rb):
        return None

    def total(f):
        return ra * epsilon_for_ratio(f / ka) + rb * epsilon_for_ratio(f / kb)

    low, high = 0.0, max(ka, kb)
    while total(high) < overlap:
        high *= 2
    for _ in range(400):
        mid = 0.5 * (low + high)
        if total(mid) < overlap:
            low = mid
        else:
            high = mid
    f = 0.0 if overlap == 0 else 0.5 * (low + high)
    return PairContact(f, epsilon_for_ratio(f / ka), epsilon_for_ratio(f / kb))


@dataclass
class WallScene:
    """Bubble moving perpendicular to a wall (:90-177): distance d from the wall surface,
    velocity positive away from it, leapfrog with h = min(rest of frame, 1/omega, d / (4|v|))."""
    radius: float
    mass: float
    stiffness: float
    damping: float
    start_distance: float
    velocity: float
    press_force: float = 0.0
    band_stiffness: float = 0.0
    frame_step: float = 1.0 / 60
    maximum_substeps: int = 1024
    frames: list = field(default_factory=list)
    substeps: list = field(default_factory=list)
    limit_exceeded: bool = False
    penetrated: bool = False

    def contact_force(self, d):
        return force(1 - d / self.radius, self.stiffness)

    def total_energy(self, d, v):
        return 0.5 * self.mass * v * v + energy(1 - d / self.radius, self.stiffness, self.radius) + self.press_force * d

    def _acceleration(self, d, v):
        return (self.contact_force(d) - self.press_force) / self.mass - self.damping * v

    def _limit(self, d, v):
        bound = self.frame_step
        k = self.band_stiffness
        eps = 1 - d / self.radius
        if eps > 0:
            k += self.stiffness * (1 + eps) / ((1 - eps) ** 3 * self.radius)
        if k > 0:
            bound = min(bound, 1 / math.sqrt(k / self.mass))
        if v != 0:
            bound = min(bound, d / (4 * abs(v)))
        return bound

    def run(self, frames: int):
        s, v = self.start_distance, self.velocity
        for _ in range(frames):
            remaining, count = self.frame_step, 0
            while remaining > 0:
                if count == self.maximum_substeps:
                    self.limit_exceeded = True
                    break
                limit = self._limit(s, v)
                h = min(remaining, limit)
                remaining = 0 if h >= remaining else remaining - h
                a0 = self._acceleration(s, v)
                half = v + 0.5 * h * a0
                s += h * half
                if s <= 0:
                    self.penetrated = True
                    count += 1
                    break
                v = half + 0.5 * h * self._acceleration(s, half)
                self.substeps.append((s, v, half, h, limit))
                count += 1
            self.frames.append((count, s, v))
            if self.limit_exceeded or self.penetrated:
                break
        return self
