"""The pins hold the oracle: each test here plants one plausible mistake in a
copy of oracle/dcm_oracle.c (a wrong density index, a wrong neighbour
index, a wrong constant), builds that copy, and runs the pin tests against
it in a subprocess; at least one pin must fail.  A control build of the
unchanged source must pass the same pins, which shows the subprocess really
loads the copy."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "dcm_oracle.c")
HDR = os.path.join(ROOT, "oracle", "dcm_oracle.h")
CFLAGS = ["-O2", "-std=c11", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-shared"]

PINS = {
    "mpdata": ["tests/test_oracle_weighted.py::test_mpdata3d_matches_generalised_textbook",
               "tests/test_oracle_weighted.py::test_mpdata_vec3d_matches_generalised_textbook"],
    "vertical": ["tests/test_oracle_weighted.py::test_vertical_mpdata_matches_generalised_textbook"],
    "sources": ["tests/test_oracle_weighted.py::test_sources_advection_of_linear_flows_closed_form"],
    "subgrid": ["tests/test_oracle_weighted.py::test_stability_functions_closed_form_values"],
}

# (name, pins, text in dcm_oracle.c, replacement)
MUTATIONS = [
    # antidiffusive: the w-face Jacobian g3 = mean of rhoan(k), rhoan(k+1)
    ("g3_one_level", "mpdata", "const double g3 = 0.5 * (rn[k] + rn[k + 1]);",
     "const double g3 = rn[k];"),
    # limiter inflow denominator: rhoa of the lower w face
    ("limiter_rhoa_index", "mpdata",
     "((pos0(A3(wt, k - 1, j, i)) * ra[k - 1]) * A3(po, k - 1, j, i)",
     "((pos0(A3(wt, k - 1, j, i)) * ra[k]) * A3(po, k - 1, j, i)"),
    # vertical antidiffusive velocity: its Jacobian gb3
    ("gb3_one_level", "vertical", "const double gb3 = 0.5 * (rn[k] + rn[k + 1]);",
     "const double gb3 = rn[k + 1];"),
    # sources: su's y term takes v of the column i+1, not i-1
    ("su_y_column", "sources",
     "A3(u, k, j - 1, i) * (A3(v, k, j - 1, i) + A3(v, k, j - 1, i + 1))",
     "A3(u, k, j - 1, i) * (A3(v, k, j - 1, i) + A3(v, k, j - 1, i - 1))"),
    # subgrid constants c and g of the stability functions
    ("subc_12", "subgrid", "SG_SUBC = 16.0", "SG_SUBC = 12.0"),
    ("subg_1", "subgrid", "SG_SUBG = 1.2", "SG_SUBG = 1.0"),
]


def _build(tmp, name, old=None, new=None):
    text = open(SRC).read()
    if old is not None:
        assert text.count(old) == 1, f"mutation {name}: pattern must occur once"
        text = text.replace(old, new)
    d = os.path.join(tmp, name)
    os.makedirs(d, exist_ok=True)
    shutil.copy(HDR, d)
    c = os.path.join(d, "dcm_oracle.c")
    open(c, "w").write(text)
    lib = os.path.join(d, "libdcm_oracle.so")
    subprocess.run(["gcc", *CFLAGS, "-o", lib, c, "-lm"], check=True)
    return lib


def _pins(lib, names):
    env = dict(os.environ, DCM_ORACLE_LIB=lib)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        *names], cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    return r.returncode, r.stdout[-2000:]


@pytest.fixture(scope="module")
def workdir(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc is needed to build the mutated oracle copies")
    return str(tmp_path_factory.mktemp("mutants"))


def test_control_copy_passes_the_pins(workdir):
    lib = _build(workdir, "control")
    names = [n for group in PINS.values() for n in group]
    rc, out = _pins(lib, names)
    assert rc == 0, out


@pytest.mark.parametrize("name,group,old,new", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_mutation_fails_a_pin(workdir, name, group, old, new):
    lib = _build(workdir, name, old, new)
    rc, out = _pins(lib, PINS[group])
    assert rc == 1, f"mutation {name} passed every pin:\n{out}"
