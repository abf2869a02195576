"""Multi-block meshes on the CPU: the reference's grid, boundary and
inner-face files read through the C ABI, the inner-face registration
(Mesh.cpp:34-187) of the library against the oracle's, and the pin of the
oracle's inner-face exchange (DataCommunication.cuh:58-110) and inner ghost
grid (Mesh.cpp:946-994): a single-block case cut into two blocks joined by
an inner face must give the single block's solution, which any wrong index,
orientation or ghost layer breaks."""
import os

import numpy as np
import pytest

import ostc
import pyoracle as po
from ostc import cases
from helpers import oracle_mesh, oracle_zone, relerr

REF = "/root/reference/input_files"
FIELDS = ("rs", "re", "face", "dir", "ts", "te", "tblk", "tface", "tdir", "src_tar", "loop", "tloop", "n")


def _flat(d):
    out = []
    for k in FIELDS:
        v = d[k]
        out.extend(v if isinstance(v, list) else [v])
    return out


def _split(axis, rotate, shape=(40, 20, 1)):
    mx, my, mz = shape
    case = cases.c1_case(mx, my) if mz == 1 else cases.c2_case(mx, my, mz)
    return case, cases.split_case(case, axis=axis, at=(mx if axis == 0 else my) * 2 // 5, rotate=rotate)


@pytest.mark.parametrize("axis,rotate,shape", [(0, False, (40, 20, 1)), (0, True, (40, 20, 1)),
                                               (1, False, (30, 24, 1)), (1, True, (30, 24, 1)),
                                               (0, True, (24, 14, 10)), (1, False, (18, 20, 10))])
def test_library_registers_inner_faces_like_the_oracle(tmp_path, axis, rotate, shape):
    case, mc = _split(axis, rotate, shape)
    root = str(tmp_path / "mb")
    cases.write_mesh_case_dir(mc, root)
    s = ostc.Solver(root, device=-1)
    s.read_mesh()
    dim = 2 if shape[2] == 1 else 3
    for b in range(2):
        got = s.inner_faces(b)
        assert got.shape[0] == 1
        src, tar = [(a, t) for (fb, a, t) in mc["inner"] if fb == b][0]
        ref = po.inner_face_register(src, tar, case["ngg"], dim)
        assert list(got[0][:29]) == _flat(ref)
    # the grid file round trip is exact (17 significant digits)
    jac, met = s.geometry(0)
    blk = mc["blocks"][0]
    jo = oracle_zone(dict(case, mx=blk["mx"], my=blk["my"], mz=blk["mz"], x=blk["x"], y=blk["y"], z=blk["z"],
                          bcs=blk["bcs"])).field("jac")
    np.testing.assert_array_equal(jac, jo)
    s.close()


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not present")
def test_reference_sample_mesh_files(tmp_path):
    """The reference's own two-block sample (grid, boundary and inner-face
    files): block extents, boundary count and the registered inner faces."""
    root = tmp_path / "ref"
    (root / "input_files" / "setup").mkdir(parents=True)
    (root / "input_files" / "setup" / "default_settings.txt").write_text(
        "int dimension = 2\nint ngg = 2\nreal gridScale = 0.001\nint species = 0\n")
    for sub in ("grid", "boundary_condition"):
        os.symlink(os.path.join(REF, sub), root / "input_files" / sub)
    s = ostc.Solver(str(root), device=-1)
    s.read_mesh()
    faces = [s.inner_faces(b) for b in range(2)]
    lines = open(os.path.join(REF, "boundary_condition", "inner   0.txt")).read().split("\n")
    pairs = [[int(v) for v in lines[k].split()] for k in (3, 4, 5, 6)]
    for b, (src, tar) in enumerate(((pairs[0], pairs[1]), (pairs[2], pairs[3]))):
        ref = po.inner_face_register(src, tar, 2, 2)
        assert list(faces[b][0][:29]) == _flat(ref)
    # block 1 (201 x 100) meets block 2 (201 x 163) along its i-max face
    assert faces[0][0][6] == 0 and faces[0][0][7] == 1 and faces[0][0][14] == 1
    jac, _ = s.geometry(0)
    assert np.all(jac[jac != 0] > 0)
    s.close()


def _mesh_vs_single(case, mc, nsteps):
    bv, sv = cases.initial_fields(case)
    z = oracle_zone(case, bv=bv, sv=sv)
    M = oracle_mesh(mc, cases.split_fields(mc, bv, sv))
    dt_z, res_z = z.run(nsteps)
    dt_m, res_m = M.run(nsteps)
    errs = []
    for b in range(2):
        got = cases.block_interior_of_full(mc, b, z.field("bv"), 6)
        g = mc["ngg"]
        blk = mc["blocks"][b]
        mine = M.zones[b].field("bv").reshape((6,) + cases.ghost_shape(blk["mx"], blk["my"], blk["mz"], g))
        mine = mine[:, g:g + blk["mz"], g:g + blk["my"], g:g + blk["mx"]]
        errs.append(relerr(mine, got))
    return max(errs), relerr(dt_m, dt_z)


@pytest.mark.parametrize("axis,rotate", [(0, False), (0, True), (1, False), (1, True)])
def test_split_mesh_reproduces_single_block_2d(axis, rotate):
    case, mc = _split(axis, rotate, (40, 24, 1))
    e, e_dt = _mesh_vs_single(case, mc, 10)
    print(f"split axis {axis} rotate {rotate}: bv {e:.2e}, dt {e_dt:.2e}")
    assert e < 1e-13 and e_dt < 1e-13


@pytest.mark.parametrize("rotate", [False, True])
def test_split_mesh_reproduces_single_block_3d(rotate):
    """3-D inviscid MUSCL/AUSM+ on the curvilinear grid, cut along i."""
    case = cases.c2_case(20, 12, 10)
    case["scheme"] = dict(case["scheme"], reconstruction=2, viscous_order=0, inviscid_scheme=3)
    case["setup"] = dict(case["setup"], **{"int reconstruction": 2, "int viscous_order": 0,
                                          "int inviscid_scheme": 3})
    case["ngg"] = 2
    case["setup"]["int ngg"] = 2
    mc = cases.split_case(case, axis=0, at=8, rotate=rotate)
    e, e_dt = _mesh_vs_single(case, mc, 5)
    print(f"3-D split rotate {rotate}: bv {e:.2e}, dt {e_dt:.2e}")
    assert e < 1e-13 and e_dt < 1e-13


def test_oracle_refuses_mixed_direction_2d_faces():
    assert po.inner_face_register([5, 5, 1, 10, 1, 1, 1], [1, 10, 1, 1, 1, 1, 2], 2, 2) is None
