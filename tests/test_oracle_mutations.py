"""The pins in test_oracle_pins.py catch the mistakes they are meant to catch.

Each case copies oracle/icf_oracle.c into a temporary directory, plants one plausible
mistake by a text substitution, builds it with the oracle's own flags and runs the pin that
must catch it against the mutant (oracle.swapped). The unmutated oracle passes the same pins
(test_oracle_pins.py). Nothing here touches the device library.
"""
import os
import shutil
import subprocess

import pytest

import test_oracle_pins as pins

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_DIR = os.path.join(os.path.dirname(HERE), "oracle")

# (name, original text, mutated text, pin that must fail)
MUTANTS = [
    ("gauss_weight_dropped", "const double W = 1.0 / 3.0;", "const double W = 1.0;",
     "test_cz_traction_resultant"),
    ("cz_area_doubled", "BtT[i] = s * -(W * area);", "BtT[i] = s * -(W * area * 2);",
     "test_cz_traction_resultant"),
    ("cz_faces_exchanged", "BtT[i] = s * -(W * area);", "BtT[i] = s * (W * area);",
     "test_cz_traction_resultant"),
    ("damping_swapped",
     "rhs[i] -= (Mij * V * C.dampingMass + Df[i * 12 + j] * C.dampingStiffness) * vn[j]",
     "rhs[i] -= (Mij * V * C.dampingStiffness + Df[i * 12 + j] * C.dampingMass) * vn[j]",
     "test_free_tet_resultants"),
    ("damping_swapped_in_tangent_too",
     "double massCoeff = V * (1.0 / (h * h) + C.dampingMass * C.NewmarkGamma / h) / C.NewmarkBeta;\n"
     "        double stiffCoeff = 1.0 + C.dampingStiffness * C.NewmarkGamma / (h * C.NewmarkBeta);",
     "double massCoeff = V * (1.0 / (h * h) + C.dampingStiffness * C.NewmarkGamma / h) / C.NewmarkBeta;\n"
     "        double stiffCoeff = 1.0 + C.dampingMass * C.NewmarkGamma / (h * C.NewmarkBeta);",
     "test_newmark_tangent_is_minus_residual_derivative"),
    ("gravity_sign_flipped", "rhs[2] += gz; rhs[5] += gz; rhs[8] += gz; rhs[11] += gz;",
     "rhs[2] -= gz; rhs[5] -= gz; rhs[8] -= gz; rhs[11] -= gz;", "test_free_tet_resultants"),
    ("beta_dropped_from_mass",
     "double massCoeff = V * (1.0 / (h * h) + C.dampingMass * C.NewmarkGamma / h) / C.NewmarkBeta;",
     "double massCoeff = V * (1.0 / (h * h) + C.dampingMass * C.NewmarkGamma / h);",
     "test_newmark_tangent_is_minus_residual_derivative"),
    ("beta_dropped_from_stiffness_damping",
     "double stiffCoeff = 1.0 + C.dampingStiffness * C.NewmarkGamma / (h * C.NewmarkBeta);",
     "double stiffCoeff = 1.0 + C.dampingStiffness * C.NewmarkGamma / h;",
     "test_newmark_tangent_is_minus_residual_derivative"),
    ("lumped_mass", "double Mij = C.M[i * 12 + j];", "double Mij = (i == j) ? C.rho / 4 : 0.0;",
     "test_consistent_mass_by_quadrature"),
    ("inertia_sign_flipped", "+ (Mij * V * an[j]);", "- (Mij * V * an[j]);", "test_free_tet_resultants"),
]


@pytest.fixture(scope="module")
def mutant_dir(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    return tmp_path_factory.mktemp("oracle_mutants")


def _build(d, name, old, new):
    src = open(os.path.join(ORACLE_DIR, "icf_oracle.c")).read()
    assert src.count(old) == 1, f"mutation target of {name} not found exactly once"
    c = d / f"m_{name}.c"
    c.write_text(src.replace(old, new))
    shutil.copy(os.path.join(ORACLE_DIR, "icf_oracle.h"), d / "icf_oracle.h")
    so = d / f"liboracle_{name}.so"
    cmd = ["gcc", "-O2", "-std=c11", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-shared",
           "-I", str(d), "-o", str(so), str(c), "-lm"]
    subprocess.run(cmd, check=True, capture_output=True)
    return str(so)


@pytest.mark.parametrize("name,old,new,pin", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_pin_catches_mutant(orc, lib, specimen_consts, mutant_dir, name, old, new, pin):
    from icflow_amd import _lib
    so = _build(mutant_dir, name, old, new)
    Cd = _lib.consts_from_prms(**{**_lib.SPECIMEN_PRMS, "dampingMass": 4.0e5, "dampingStiffness": 7.0e-7})
    fn = getattr(pins, pin)
    with orc.swapped(so):
        orc.set_constants(specimen_consts)
        args = {"test_cz_traction_resultant": (orc, specimen_consts),
                "test_consistent_mass_by_quadrature": (orc, specimen_consts)}.get(pin, (orc, Cd))
        with pytest.raises(AssertionError):
            fn(*args)
    orc.set_constants(specimen_consts)
