"""The oracle against what the mathematics fixes (tests/oracle_pins.py), and
each pin against a deliberately broken copy of the oracle: a mistake
planted in the oracle's source (a downwind flux, another SSP stage weight,
a first-order column reference, a flipped terrain correction, a linear
taper) must make its pin fail, so the pins are shown to pin something.
CPU only; the mutants are compiled from oracle/gwm_oracle.cpp into a
temporary directory."""
import concurrent.futures
import os
import shutil
import subprocess
import sys

import pytest

import oracle_pins

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE = os.path.join(os.path.dirname(HERE), "oracle")


@pytest.mark.parametrize("name", list(oracle_pins.PINS))
def test_pin_holds_on_oracle(orc, name):
    oracle_pins.PINS[name](orc)


# (pin that must catch it, planted mistake: text of the oracle and its replacement)
MUTANTS = {
    "tracer_downwind": ("tracer_ssp_shift",
                        "const R fw = mw * (mw >= R(0) ? qw : qc);",
                        "const R fw = mw * (mw >= R(0) ? qc : qw);"),
    "tracer_stage2_weights": ("tracer_ssp_shift",
                              "bl(q2[n], q0[n], tmp[n], R(0.75), R(0.25));",
                              "bl(q2[n], q0[n], tmp[n], R(0.5), R(0.5));"),
    "dry_stage3_weights": ("dry_ssp_shift",
                           "blend(S[n], q0[n], tmp[n], R(1) / R(3), R(2) / R(3));",
                           "blend(S[n], q0[n], tmp[n], R(0.5), R(0.5));"),
    "theta_downwind": ("theta_flux_shift",
                       "const R fe = me * (me >= R(0) ? thc : the);",
                       "const R fe = me * (me >= R(0) ? the : thc);"),
    "momentum_downwind": ("dry_ssp_shift",
                          "return a >= R(0) ? a * ql : a * qr;",
                          "return a >= R(0) ? a * qr : a * ql;"),
    "w_momentum_downwind": ("w_momentum_shift",
                            "return a >= R(0) ? a * ql : a * qr;",
                            "return a >= R(0) ? a * qr : a * ql;"),
    "self_advection_downwind": ("u_self_advection_diffuses",
                                "const R fw = upw(uc(i - 1, j, k), qw, qc);",
                                "const R fw = upw(uc(i - 1, j, k), qc, qw);"),
    "column_reference_explicit": ("column_reference_second_order",
                                  "pn = pref[k - 1] - g * R(0.5) * (rref[k - 1] + rn) * dz;",
                                  "pn = pref[k - 1] - g * rref[k - 1] * dz;"),
    "terrain_correction_sign": ("terrain_pgf_cancels",
                                "return (pr - pl) / dx - dzdx * dpdz;",
                                "return (pr - pl) / dx + dzdx * dpdz;"),
    "taper_linear": ("taper_smoothstep",
                     "  if (eta <= taper) return 1.0;\n  const double s = std::clamp((eta - taper) / "
                     "std::max(1.0e-9, 1.0 - taper), 0.0, 1.0);\n  return 1.0 - s * s * (3.0 - 2.0 * s);",
                     "  if (eta <= taper) return 1.0;\n  const double s = std::clamp((eta - taper) / "
                     "std::max(1.0e-9, 1.0 - taper), 0.0, 1.0);\n  return 1.0 - s;"),
}


def _build(name, tmp):
    _, old, new = MUTANTS[name]
    src = open(os.path.join(ORACLE, "gwm_oracle.cpp")).read()
    assert src.count(old) >= 1, f"mutation site of {name} not found"
    d = os.path.join(tmp, name)
    os.makedirs(d, exist_ok=True)
    shutil.copy(os.path.join(ORACLE, "gwm_oracle.h"), d)
    with open(os.path.join(d, "gwm_oracle.cpp"), "w") as fh:
        fh.write(src.replace(old, new))
    so = os.path.join(d, "liboracle.so")
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-fopenmp", "-shared", "-o", so,
                    os.path.join(d, "gwm_oracle.cpp")], check=True)
    return so


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    tmp = str(tmp_path_factory.mktemp("oracle_mutants"))
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 2)) as ex:
        futs = {n: ex.submit(_build, n, tmp) for n in MUTANTS}
        return {n: f.result() for n, f in futs.items()}


@pytest.mark.parametrize("name", list(MUTANTS))
def test_pin_catches_planted_mistake(mutant_libs, name):
    pin = MUTANTS[name][0]
    code = ("import sys; sys.path[:0] = [%r, %r]; import pyoracle, oracle_pins; "
            "oracle_pins.PINS[%r](pyoracle)" % (ORACLE, HERE, pin))
    env = dict(os.environ, GWM_ORACLE_LIB=mutant_libs[name], OMP_NUM_THREADS="2")
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode != 0, f"pin {pin} passed on the mutant {name}"
    assert "AssertionError" in r.stderr, r.stderr[-2000:]
