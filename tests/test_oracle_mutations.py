"""The oracle's pins must catch plausible misreadings of the reference.

Each case plants one mistake in a copy of oracle/cf_oracle.py (a text edit of
the source, loaded as a fresh module on the CPU) and requires at least one of
the physics pins in tests/oracle_pins.py to fail on it.  The mistakes are the
kind a restatement of PF:1438-1540 (Lamb vector, h_v, h_g), PF:3348-3534
(operators and right-hand sides), PF:401-474 (mean mode) and PF:2347-2407
(kinematics) could contain: a flipped sign, a swapped product pair, a wrong
coefficient, a shifted row, the wrong time scheme.
"""
import os
import sys
import types

import pytest

from . import oracle_pins

ORACLE_SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle",
                          "cf_oracle.py")

# (name, original text, planted text, expected number of occurrences)
MUTATIONS = [
    ("h_g sign", "hg = 1j * a3 * lamb3 - 1j * b3 * lamb1", "hg = -(1j * a3 * lamb3 - 1j * b3 * lamb1)", 1),
    ("i beta D L_z sign", "hv = hv + 1j * b3 * dlz", "hv = hv - 1j * b3 * dlz", 1),
    ("i alpha D L_x sign", "hv = k2 * lamb2 + 1j * a3 * dlx", "hv = k2 * lamb2 - 1j * a3 * dlx", 1),
    ("k^2 L_y sign", "hv = k2 * lamb2 + 1j * a3 * dlx", "hv = -k2 * lamb2 + 1j * a3 * dlx", 1),
    ("L_z product pair swapped", "lamb3 = getlamb(ox, v, p) - getlamb(oy, u, p)",
     "lamb3 = getlamb(oy, u, p) - getlamb(ox, v, p)", 1),
    ("L_x product pair swapped", "lamb1 = getlamb(oy, w, p, shift=p.re_p) - getlamb(oz, v, p)",
     "lamb1 = getlamb(oz, v, p, shift=p.re_p) - getlamb(oy, w, p)", 1),
    ("L_y product pair swapped", "lamb2 = getlamb(oz, u, p) - getlamb(ox, w, p)",
     "lamb2 = getlamb(ox, w, p) - getlamb(oz, u, p)", 1),
    ("re_p shift sign", "shift=p.re_p)", "shift=-p.re_p)", 1),
    ("omega_y operator k^2 coefficient dt/re",
     "(1 + kmn * p.dt * 0.5 / p.re) * self.t0[rows]", "(1 + kmn * p.dt / p.re) * self.t0[rows]", 1),
    ("omega_y explicit operator k^2 coefficient dt/re",
     "p.re + (1 - kmn * p.dt * 0.5 / p.re) * self.t0[rows]", "p.re + (1 - kmn * p.dt / p.re) * self.t0[rows]", 1),
    ("v operator 2ck^2 halved", "(1 + kmn * p.dt / p.re) * t2[rows]", "(1 + kmn * p.dt * 0.5 / p.re) * t2[rows]",
     1),
    ("v explicit operator 2ck^2 halved", "(1 - kmn * p.dt / p.re) * self.t2[rows]",
     "(1 - kmn * p.dt * 0.5 / p.re) * self.t2[rows]", 1),
    ("v operator fourth-derivative sign", "a[4:] = (-t4[rows]", "a[4:] = (t4[rows]", 1),
    ("mean-mode operator diffusion sign", "a[2:] = -self.t2[rows] * p.dt * 0.5 / p.re + self.t0[rows]",
     "a[2:] = self.t2[rows] * p.dt * 0.5 / p.re + self.t0[rows]", 1),
    ("v forcing AB2 -> Euler", "f = (3.0 * dt / 2.0 * hv[:, jj, ii] - dt / 2.0 * u[:, jj, ii]).T",
     "f = (dt * hv[:, jj, ii]).T", 1),
    ("v forcing rows shifted", "rhs[:, 4:] = np.einsum(\"mpl,ml->mp\", bv, vm) + f[:, 2:n - 1]",
     "rhs[:, 4:] = np.einsum(\"mpl,ml->mp\", bv, vm) + f[:, 3:n]", 1),
    ("omega_y forcing rows shifted", "rhs[:, 2:] = np.einsum(\"mpl,ml->mp\", be, om) + f[:, 1:n]",
     "rhs[:, 2:] = np.einsum(\"mpl,ml->mp\", be, om) + f[:, 0:n - 1]", 1),
    ("mean forcing sign", "f = -1.0 * dt / 2.0 * (3.0 * st[cn] - st[cnm1])",
     "f = 1.0 * dt / 2.0 * (3.0 * st[cn] - st[cnm1])", 1),
    ("mean omega_x = -DW after the solve", "(\"clamb3n\", \"clamb3nm1\", \"w0\", \"wnp1\", \"omegaxnp1\", 1.0)",
     "(\"clamb3n\", \"clamb3nm1\", \"w0\", \"wnp1\", \"omegaxnp1\", -1.0)", 1),
    ("mean omega_x = -DW in the kinematics", "st[\"omegaxnp1\"][:, 0, 0] = dz(st[\"w0\"])",
     "st[\"omegaxnp1\"][:, 0, 0] = -dz(st[\"w0\"])", 1),
    ("kinematics u: i beta omega_y sign", "uu = (1j * a_ * dv - 1j * b_ * on) / k2s",
     "uu = (1j * a_ * dv + 1j * b_ * on) / k2s", 1),
    ("solve kinematics omega_z sign", "oz[:, jj, ii] = -dz(uu) + 1j * a_ * vt", "oz[:, jj, ii] = dz(uu) + 1j * a_ * vt",
     1),
    ("Chebyshev analysis end-point weight", "res[0] /= 2.0", "res[0] /= 1.0", 1),
]


def load_mutant(original, planted, count):
    with open(ORACLE_SRC) as f:
        src = f.read()
    assert src.count(original) == count, f"mutation anchor not found as expected: {original!r}"
    name = "cf_oracle_mutant"
    mod = types.ModuleType(name)
    sys.modules[name] = mod          # dataclasses resolve annotations through sys.modules
    try:
        exec(compile(src.replace(original, planted), ORACLE_SRC + " (mutant)", "exec"), mod.__dict__)
    finally:
        del sys.modules[name]
    return mod


@pytest.mark.parametrize("name,original,planted,count", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_pins_catch_planted_mistake(name, original, planted, count):
    mutant = load_mutant(original, planted, count)
    caught = []
    for pin in oracle_pins.ALL_PINS:
        try:
            pin(mutant)
        except (AssertionError, FloatingPointError, ValueError, ZeroDivisionError) as e:
            caught.append(f"{pin.__name__}: {str(e)[:80]}")
    assert caught, f"no pin caught the planted mistake: {name}"
    print(f"{name}: caught by {'; '.join(caught)}")


def test_unmutated_copy_passes_every_pin():
    mod = load_mutant("def dz(", "def dz(", 1)       # an identity edit: the loader itself changes nothing
    for pin in oracle_pins.ALL_PINS:
        pin(mod)
