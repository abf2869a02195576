"""Host side of the C ABI without a GPU: the library loads and exports every
entry point of include/ostc.h; setup-file, CHEMKIN/therm/tran readers and
grid metrics (host-only solver, device = -1) against the reference's own
parser output and the oracle."""
import os
import re

import numpy as np
import pytest

import ostc
import pyoracle as po
from ostc import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "ostc.h")).read()
    return sorted(set(re.findall(r"\b(ostc_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = ostc.lib()
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing
    assert L.ostc_abi_version() == 1


def _case_dir(tmp_path, species=1, reaction=1):
    case = cases.c1_case(20, 12)
    case["setup"]["int species"] = species
    case["setup"]["int reaction"] = reaction
    root = str(tmp_path / "case")
    cases.write_case_dir(case, root)
    return root


def test_mechanism_tables_match_reference_parser(tmp_path, tables):
    s = ostc.Solver(_case_dir(tmp_path), device=-1)
    ns, nr = s.get_int("n_spec"), s.get_int("n_reac")
    assert (ns, nr) == (tables["n_spec"], tables["n_reac"]) == (9, 19)
    for key in ("mw", "t_low", "t_mid", "t_high", "LJ_potent_inv", "vis_coeff", "A", "b", "Ea", "A2", "b2", "Ea2",
                "troe_alpha", "troe_t3", "troe_t1", "troe_t2", "third_body_coeff", "WjDivWi_to_One4th",
                "sqrt_WiDivWjPl1Mul8", "binary_diffusivity_coeff", "kb_over_eps_jk", "label", "stoi_f", "stoi_b",
                "order"):
        ref = np.asarray(tables[key], dtype=np.float64)
        got = s.table(key, ref.size)
        np.testing.assert_array_equal(got, ref, err_msg=key)
    for key, name in (("high_coeff", "high_coeff"), ("low_coeff", "low_coeff")):
        ref = np.asarray(tables[key])
        np.testing.assert_array_equal(s.table(name, ref.size), ref)
    s.close()


def test_mixing_mechanism_matches_reference_parser(tmp_path):
    """The reference's 4-species mixing mechanism (H2, H2O, O2, N2) read
    without reactions: the species tables equal the reference parser's
    (tests/golden/h2_mixing_4sp_tables.txt) bit for bit."""
    tables = po.load_tables(os.path.join(ROOT, "tests", "golden", "h2_mixing_4sp_tables.txt"))
    case = cases.mixing4_case(24, 12)
    root = str(tmp_path / "m4")
    cases.write_case_dir(case, root)
    s = ostc.Solver(root, device=-1)
    assert (s.get_int("n_spec"), s.get_int("n_reac")) == (tables["n_spec"], tables["n_reac"]) == (4, 0)
    assert s.get_int("n_var") == 9
    for key in ("mw", "t_low", "t_mid", "t_high", "LJ_potent_inv", "vis_coeff", "high_coeff", "low_coeff",
                "WjDivWi_to_One4th", "sqrt_WiDivWjPl1Mul8", "binary_diffusivity_coeff", "kb_over_eps_jk"):
        ref = np.asarray(tables[key], dtype=np.float64)
        np.testing.assert_array_equal(s.table(key, ref.size), ref, err_msg=key)
    s.close()


def test_uncompiled_species_count_is_refused(tmp_path):
    """A mechanism whose species count has no compiled kernels fails at
    ostc_init with a message naming the compiled counts."""
    case = cases.mixing4_case(24, 12)
    root = str(tmp_path / "m5")
    cases.write_case_dir(case, root)
    mech = os.path.join(root, "input_files", "chemistry", "five.inp")
    with open(mech, "w") as f:
        f.write("ELEMENTS\nH O N\nEND\nSPECIES\nH2 H2O O2 N2 OH\nEND\nREACTIONS\nEND\n")
    setup = os.path.join(root, "input_files", "setup", "default_settings.txt")
    txt = open(setup).read().replace("chemistry/H2_mech_mixing.inp", "chemistry/five.inp")
    open(setup, "w").write(txt)
    with pytest.raises(ostc.OstcError, match="compiled for 4 and 9 species"):
        ostc.Solver(root, device=-1)


def test_setup_derivations(tmp_path):
    s = ostc.Solver(_case_dir(tmp_path, species=0, reaction=0), device=-1)
    assert s.get_int("ngg") == 2
    assert s.get_int("inviscid_type") == 0
    assert s.get_int("n_var") == 5
    assert s.get_real("cfl") == 0.5
    s.close()
    case = cases.c2_case(16, 12, 10)
    root = str(tmp_path / "c2")
    cases.write_case_dir(case, root)
    s = ostc.Solver(root, device=-1)
    assert s.get_int("ngg") == 3
    assert s.get_int("inviscid_type") == 1  # Roe + WENO5 runs the AWENO path
    s.close()


def test_inflow_state_matches_reference_construction(tmp_path):
    case = cases.c1_case(20, 12)
    root = str(tmp_path / "c1")
    cases.write_case_dir(case, root)
    s = ostc.Solver(root, device=-1)
    got = s.inflow_state("freestream")
    ref = po.inflow_state(2.0, 101325.0, 300.0)
    for i, k in enumerate(("rho", "u", "v", "w", "p", "t", "vel")):
        assert got[i] == pytest.approx(ref[k], rel=1e-15, abs=1e-300), k
    s.close()


@pytest.mark.parametrize("shape", [(24, 16, 1), (20, 14, 12)])
def test_grid_metrics_match_oracle(tmp_path, shape):
    mx, my, mz = shape
    if mz == 1:
        case = cases.c1_case(mx, my)
    else:
        case = cases.c2_case(mx, my, mz)
    root = str(tmp_path / "g")
    cases.write_case_dir(case, root)
    s = ostc.Solver(root, device=-1)
    s.add_block(mx, my, mz, case["x"], case["y"], case["z"])
    jac, met = s.geometry()
    g = s.get_int("ngg")
    n = (mx + 2 * g) * (my + 2 * g) * (mz + 2 * g)
    nc = (mx + 2 * g + 2) * (my + 2 * g + 2) * (mz + 2 * g + 2)
    xg, yg, zg = np.zeros(nc), np.zeros(nc), np.zeros(nc)
    jo, mo = np.zeros(n), np.zeros(9 * n)
    po.lib().orc_grid_metrics(mx, my, mz, g, *(po._p(np.ascontiguousarray(a, dtype=np.float64).ravel())
                                                 for a in (case["x"], case["y"], case["z"])),
                              po._p(xg), po._p(yg), po._p(zg), po._p(jo), po._p(mo))
    np.testing.assert_array_equal(jac, jo)
    np.testing.assert_array_equal(met, mo.reshape(-1, 9).T.ravel())
    s.close()
