"""Edge cases of the inputs against the FP64 oracle: the smallest mesh the
solver takes (2 x 2 elements, where the top-row Lysmer strip of 3*hpn nodes
is longer than the row and the one tile holds two element columns),
receivers on the first and last node columns, meshes a whole number of
storage blocks wide, a run of a single level
(nst = 1: only GRT_Interface's level 0 is recorded), and a mesh too large
for 32-bit node offsets."""
import os

import numpy as np
import pytest

import pyoracle
import semsks
from semsks import inputs

pytestmark = pytest.mark.gpu

TOL = 1e-4


def run_both(d, **eng):
    m = semsks.Model(["par=run.par"])
    with semsks.Engine(**eng) as e:
        s = e.admit(m.problem)
        e.advance(m.problem.nst - 1)
        tr = e.traces(s)
    ref = pyoracle.Oracle(["par=run.par"], "f64", cwd=d).run()
    return tr, ref


def check(tr, ref):
    assert tr.shape == ref.shape
    a, b = tr.astype(np.float64), ref.astype(np.float64)
    for k in range(3):
        nb = np.linalg.norm(b[:, k])
        assert nb > 0
        l2 = np.linalg.norm(a[:, k] - b[:, k]) / nb
        pk = np.abs(a[:, k] - b[:, k]).max() / np.abs(b[:, k]).max()
        assert l2 <= TOL and pk <= TOL, (k, l2, pk)


@pytest.mark.parametrize("tile_z", [1, 0])
def test_smallest_mesh_vs_fp64(tmp_path, monkeypatch, tile_z):
    print("case: 2x2 elements, 300 steps, receivers at ix 0 and 7, tile_z", tile_z, flush=True)
    d = os.fspath(tmp_path)
    inputs.small_case(nex=2, layers_thk=(2.0, 2.002), nst=300, nrec=2, rec_x0=0.0, rec_x1=4.0,
                      psi_deg=30.0, sh_zero=False).write(d)
    monkeypatch.chdir(d)
    tr, ref = run_both(d, tile_z=tile_z)
    assert np.abs(ref).max() > 0.5
    check(tr, ref)


def test_receivers_on_the_side_columns_vs_fp64(tmp_path, monkeypatch):
    print("case: 24x12, receivers at ix 0 and nx-2, 400 steps", flush=True)
    d = os.fspath(tmp_path)
    inputs.small_case(nst=400, nrec=2, rec_x0=0.0, rec_x1=48.0, psi_deg=30.0, sh_zero=False).write(d)
    monkeypatch.chdir(d)
    m = semsks.Model(["par=run.par"])
    ix, _ = m.receivers()
    assert ix[0] == 0 and ix[1] >= m.problem.nex * m.problem.hpn - 1
    tr, ref = run_both(d)
    check(tr, ref)


@pytest.mark.parametrize("hpn,nex", [(4, 22), (3, 15)])
def test_mesh_width_on_a_storage_block_boundary_vs_fp64(tmp_path, monkeypatch, hpn, nex):
    """Nodes are stored in column blocks of one tile's owned width (11
    elements at hpn = 4, 15 at hpn = 3).  A mesh a whole number of blocks
    wide puts its last node column alone in a block of its own.  Receivers
    on the first node column and next to the last one (whose force reads
    that lone column), against the FP64 oracle."""
    print("case: %dx12 elements, hpn %d, receivers at ix 0 and nx-2" % (nex, hpn), flush=True)
    d = os.fspath(tmp_path)
    inputs.small_case(nex=nex, hpn=hpn, nst=400, nrec=2, rec_x0=0.0, rec_x1=2.0 * nex, psi_deg=30.0,
                      sh_zero=False).write(d)
    monkeypatch.chdir(d)
    m = semsks.Model(["par=run.par"])
    ix, _ = m.receivers()
    assert ix[0] == 0 and ix[1] >= m.problem.nex * m.problem.hpn - 1
    tr, ref = run_both(d)
    assert np.abs(ref).max() > 0.1
    check(tr, ref)


def test_single_level(tmp_path, monkeypatch):
    """nst = 1: no step is taken and the one record is level 0.  The
    interpolation ring gives weight 1 to a slot not yet read at it = 0
    (interpl.c:59-91), so GRT_Interface injects zeros: the traces are zero,
    and must be so bit for bit (the -w component as -0.0)."""
    print("case: 24x12, nst = 1", flush=True)
    d = os.fspath(tmp_path)
    inputs.small_case(nst=1, psi_deg=30.0, sh_zero=False).write(d)
    monkeypatch.chdir(d)
    m = semsks.Model(["par=run.par"])
    with semsks.Engine() as e:
        s = e.admit(m.problem)
        e.advance(0)
        tr = e.traces(s)
    ref = pyoracle.Oracle(["par=run.par"], "f64", cwd=d).run()
    assert tr.shape == ref.shape == (1, 3, 8)
    assert np.array_equal(tr.view(np.uint32), ref.view(np.uint32))


def test_mesh_beyond_32bit_node_offsets_is_rejected():
    """A mesh whose level plane needs more than 2^31 nodes is turned away
    at admission with SEMSKS_EINVAL, before any of its arrays is read or
    any device memory is taken; the engine then admits a normal case."""
    print("case: admit a 60000x10000-element mesh (9.6e9 nodes)", flush=True)
    import ctypes as C
    mat = (C.c_float * semsks.MAT_STRIDE)(3.3, 200.0)
    ids = (C.c_uint16 * 4)()
    p = semsks.Problem(nex=60000, nez=10000, hpn=4, neh=1, eh=2.0, sdt=0.01, nst=10, solver=0, absorb=1,
                       n_mat=1, mat_table=mat, mat_id=ids, ntrace=0, nt=0, dp=0.01)
    with semsks.Engine() as e:
        with pytest.raises(semsks.SemsksError) as err:
            e.admit(p)
        assert err.value.code == 1
        from semsks import configs
        b = configs.Built(inputs.small_case(nst=20))
        e.admit(b.model.problem)
