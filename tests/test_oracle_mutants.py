"""Each pin of tests/test_law_pins.py, tests/test_transport_pins.py and
tests/test_microstate_pins.py is only worth the mistakes it catches.  This test plants single-line mistakes
in copies of oracle/dem_oracle.c (compiled into a temporary directory, on the
CPU), runs the pins against each copy and requires every copy to fail at
least one of them; the unmodified oracle passes them all (the other tests).

The mistakes are the ones a restatement of the reference's law most easily
makes and that a symmetric test cannot see: the end-point spin in the
half-step quaternion, a flipped quaternion derivative or one of its
cross terms, a flipped gyroscopic
term, the partner's lever arm in the pair couple, flipped cross products in
the wall and manifold rolling rates, the particle-triangle couple and the
body friction torque; history carried to a new tangent plane without its
length, the pair spring grown by the end-point rate, expiry without its
one-step lag and the unscaled yield fallback; the microstate's springs
swapped and its friction-torque lever reversed."""
import importlib
import inspect
import os
import subprocess

import pytest

from conftest import ROOT
from dem_mbd_cuda_amd._abi import Library

import test_law_pins as P

SRC = os.path.join(ROOT, "oracle", "dem_oracle.c")

# (name, original text, mistaken text, occurrence to change (0-based), pins)
MUTANTS = [
    ("half-step quaternion from the end-point spin",
     "quat_derivative(qhw, qhx, qhy, qhz, s->whx[i], s->why[i], s->whz[i],",
     "quat_derivative(qhw, qhx, qhy, qhz, s->wx[i], s->wy[i], s->wz[i],", 0,
     ["free_spinners", "free_body_invariants"]),
    ("quaternion derivative with the wrong sign",
     "    *dqw = 0.5 * (-qx * wx - qy * wy - qz * wz);\n    *dqx = 0.5 * (qw * wx + qy * wz - qz * wy);\n"
     "    *dqy = 0.5 * (qw * wy - qx * wz + qz * wx);\n    *dqz = 0.5 * (qw * wz + qx * wy - qy * wx);",
     "    *dqw = -0.5 * (-qx * wx - qy * wy - qz * wz);\n    *dqx = -0.5 * (qw * wx + qy * wz - qz * wy);\n"
     "    *dqy = -0.5 * (qw * wy - qx * wz + qz * wx);\n    *dqz = -0.5 * (qw * wz + qx * wy - qy * wx);", 0,
     ["free_spinners", "spin_closed_form"]),
    ("quaternion product in the wrong order (q x w for w x q)",
     "*dqx = 0.5 * (qw * wx + qy * wz - qz * wy);", "*dqx = 0.5 * (qw * wx - qy * wz + qz * wy);", 0,
     ["free_spinners"]),
    ("gyroscopic term with the wrong sign",
     "s->awy[i] = (ty - (ix - iz) * wx * wz) / iy;", "s->awy[i] = (ty + (ix - iz) * wx * wz) / iy;", 0,
     ["free_spinners", "euler_first_step"]),
    ("pair couple about the partner's lever arm",
     "mul(cross(dir, friction), li)", "mul(cross(dir, friction), lj)", 0, ["unequal_pair"]),
    ("wall rolling rate with the cross product reversed",
     "rollRate = add(rollRate, mul(cross(normal, tv), 0.5 / lever));",
     "rollRate = add(rollRate, mul(cross(tv, normal), 0.5 / lever));", 0, ["wall_slider"]),
    ("manifold rolling rate with the cross product reversed",
     "const V3 rollRate = add(sub(rw, twistRate), mul(cross(normal, tv), 0.5 / fmax(distance, 1.0e-14)));",
     "const V3 rollRate = add(sub(rw, twistRate), mul(cross(tv, normal), 0.5 / fmax(distance, 1.0e-14)));", 1,
     ["body_slider"]),
    ("particle-triangle couple reversed",
     "const V3 couple = mul(add(moment, mul(cross(normal, friction), distance)), scale);",
     "const V3 couple = mul(add(moment, mul(cross(friction, normal), distance)), scale);", 0, ["body_slider"]),
    ("body friction torque reversed",
     "bft = mul(cross(arm, friction), -scale)", "bft = mul(cross(friction, arm), -scale)", 0, ["body_slider"]),
    ("history carried to the new tangent plane without its length",
     "return mul(projected, pn / qn);", "return projected;", 0,
     ["test_transport_pins:test_transport_under_a_rotating_normal"]),
    ("pair spring grown by the end-point rate",
     "const V3 friction = limited_rates(mtv, etv, kS, cS,", "const V3 friction = limited_rates(etv, etv, kS, cS,", 0,
     ["test_transport_pins:test_spring_takes_the_half_step_rate_damping_the_endpoint_rate"]),
    ("history expired without the one-step lag",
     "lastSeen[index] < step - 1", "lastSeen[index] < step", 0,
     ["test_transport_pins:test_history_expires_one_step_after_contact_is_lost"]),
    ("yield with no increment keeps the stored spring unscaled",
     "response = pn > 1.0e-14 ? mul(previous, limit / pn) : v3(0.0, 0.0, 0.0);", "response = previous;", 0,
     ["test_transport_pins:test_yield_with_no_increment_rescales_the_stored_spring"]),
    ("microstate with the tangential and twisting springs swapped",
     "const V3 th = hload(s->hbv, slot, 1), wh = hload(s->hbv, slot, 2);",
     "const V3 th = hload(s->hbv, slot, 2), wh = hload(s->hbv, slot, 1);", 0,
     ["test_microstate_pins:test_sliding_on_an_off_centre_body"]),
    ("microstate friction torque with the lever-arm product reversed",
     "const V3 btt = cross(arm, bth);", "const V3 btt = cross(bth, arm);", 0,
     ["test_microstate_pins:test_sliding_on_an_off_centre_body"]),
]


def _build(tmp_path, idx, old, new, occurrence):
    text = open(SRC).read()
    parts = text.split(old)
    assert len(parts) > occurrence + 1, f"mutant {idx}: '{old}' occurs {len(parts) - 1} times"
    mutated = old.join(parts[:occurrence + 1]) + new + old.join(parts[occurrence + 1:])
    src = tmp_path / f"mutant{idx}.c"
    src.write_text(mutated)
    so = tmp_path / f"libmutant{idx}.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-fopenmp", "-ffp-contract=off", "-std=c11", "-shared", "-I", os.path.join(ROOT, "oracle"),
                    "-o", str(so), str(src), "-lm"], check=True)
    return Library(str(so), "ora_")


def _pin(name, lib, tmp_path):
    if name in P.law_cases.CASES:
        P.legacy_pin(lib, name, tmp_path)
    elif ":" in name:
        module, func = name.split(":")
        f = getattr(importlib.import_module(module), func)
        f(*([lib, tmp_path] if "tmp_path" in inspect.signature(f).parameters else [lib]))
    else:
        getattr(P, name)(lib)


@pytest.mark.parametrize("idx", range(len(MUTANTS)), ids=[m[0] for m in MUTANTS])
def test_planted_mistake_fails_a_pin(tmp_path, idx):
    name, old, new, occurrence, pins = MUTANTS[idx]
    lib = _build(tmp_path, idx, old, new, occurrence)
    caught = []
    for pin in pins:
        try:
            _pin(pin, lib, tmp_path)
        except AssertionError as e:
            caught.append((pin, str(e).splitlines()[0][:160]))
    assert caught, f"no pin catches: {name}"
    # every listed pin catches it
    assert len(caught) == len(pins), (name, caught)
