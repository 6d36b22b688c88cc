"""The closed-form checks catch plausible slips in the oracle (CPU only).

Each case plants one mistake of the kind a restatement of getTempVari_2 and
getDtF (TTSH3D.cu:561-632) could make -- a transposed index, a flipped sign, a
dropped factor, the reference's integer 1/3 read as a real third -- into a
private copy of oracle/ttsh_oracle.c, builds that copy with gcc into a
temporary directory, and runs the checks of test_oracle_coupling.py against
it.  At least one check must fail for every mutant; the unmodified source
must pass them all.  Nothing here touches the built oracle or the product.
"""
import importlib.util
import os
import subprocess

import pytest

import test_oracle_coupling as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "ttsh_oracle.c")
ORACLE_PY = os.path.join(ROOT, "oracle", "oracle.py")

# (name, text found exactly once in ttsh_oracle.c, its replacement)
MUTANTS = [
    ("sigma_transposed", "s->G[3 * b + a]", "s->G[3 * a + b]"),
    ("active_stress_sign", "+ P.cA * (d1x(s->Sig[0]", "- P.cA * (d1x(s->Sig[0]"),
    ("capillary_sign", "+ P.mu * lap(s->vx, nb, h) + mu0 * ax", "+ P.mu * lap(s->vx, nb, h) - mu0 * ax"),
    ("capillary_wrong_axis", "+ P.mu * lap(s->vy, nb, h) + mu0 * ay", "+ P.mu * lap(s->vy, nb, h) + mu0 * ax"),
    ("lambda0_sign", "- P.lambda0 * (px * dxpx + py * dypx + pz * dzpx)",
     "+ P.lambda0 * (px * dxpx + py * dypx + pz * dzpx)"),
    ("lambda0_transposed", "- P.lambda0 * (px * dxpx + py * dypx + pz * dzpx)",
     "- P.lambda0 * (px * dxpx + py * dxpy + pz * dxpz)"),
    ("phi_advection_sign", "s->dtphi[idx] = -vx * ax", "s->dtphi[idx] = vx * ax"),
    ("v_advection_transposed", "- vy * d1y(s->vx, nb, h)", "- vy * d1x(s->vy, nb, h)"),
    ("cahn_hilliard_coefficient", "(phi - 3 * phi * phi + 2 * phi * phi * phi)",
     "(phi - 2 * phi * phi + 2 * phi * phi * phi)"),
    ("gamma_on_phi", "P.gamma * (lap(s->mu0, nb, h))", "P.gamma * (lap(s->phi, nb, h))"),
    ("molecular_field_without_phi", "s->mux[idx] = phi * px +", "s->mux[idx] = px +"),
    ("alpha_without_phi", "- P.alpha * phi * px", "- P.alpha * px"),
    ("anchoring_P2_for_modulus", "+ sP2 * (s->Axx[idx] * px", "+ P2 * (s->Axx[idx] * px"),
    ("anchoring_real_third", "s->Axx[idx] = ax * ax - 1 / 3 * dP2", "s->Axx[idx] = ax * ax - 1.0 / 3 * dP2"),
    ("corotation_not_zero", "s->OPx[idx] = 1 / 2 * (", "s->OPx[idx] = 0.5 * ("),
    ("eta_sign", "+ P.eta * lap(s->mux, nb, h)", "- P.eta * lap(s->mux, nb, h)"),
]


def _load_oracle_copy(lib_path, name):
    spec = importlib.util.spec_from_file_location(f"oracle_{name}", ORACLE_PY)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    mod._LIB_PATH = lib_path  # lib() loads lazily from this path
    return mod


def _build(tmp_path, name, text):
    c = tmp_path / f"{name}.c"
    c.write_text(text)
    so = tmp_path / f"lib_{name}.so"
    subprocess.run(["gcc", "-O1", "-fno-fast-math", "-ffp-contract=off", "-fopenmp", "-fPIC", "-std=c11",
                    "-shared", "-o", str(so), str(c), "-lm"], check=True)
    return str(so)


def _failures(mod):
    failed = []
    for check in T.CHECKS:
        try:
            check(mod)
        except AssertionError:
            failed.append(check.__name__)
    return failed


def test_unmodified_copy_passes(tmp_path):
    with open(SRC) as f:
        src = f.read()
    mod = _load_oracle_copy(_build(tmp_path, "orig", src), "orig")
    assert _failures(mod) == []


@pytest.mark.parametrize("name,old,new", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_mutant_is_caught(tmp_path, name, old, new):
    with open(SRC) as f:
        src = f.read()
    assert src.count(old) == 1, f"mutation site of {name} not found exactly once"
    mod = _load_oracle_copy(_build(tmp_path, name, src.replace(old, new)), name)
    failed = _failures(mod)
    assert failed, f"no closed-form check caught the mutant {name}"
