"""semsks_stiff_dev, the K u entry point that allocates nothing (its
scratch is a fixed per-GPU device array), against semsks_stiff on the same
slot and field: bit-exact, on two calls in a row, the second on a smaller
mesh, then once more on the first; a mesh larger than the scratch is
refused with EINVAL before any device work.  The field and force buffers
are the caller's device allocations (made here with the HIP runtime the
library itself links)."""
import ctypes as C
import os

import numpy as np
import pytest

import semsks
from semsks import aniso, inputs

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    semsks.lib()                          # loads libsemsks.so and the HIP runtime it links
    H = C.CDLL("libamdhip64.so.7")        # that same, already loaded runtime
    H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    H.hipFree.argtypes = [C.c_void_p]
    H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return H


class DevBuf:
    def __init__(self, H, nbytes):
        self.H, self.p = H, C.c_void_p()
        assert H.hipMalloc(C.byref(self.p), nbytes) == 0
        self.n = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.n
        assert self.H.hipMemcpy(self.p, a.ctypes.data, self.n, H2D) == 0

    def get(self, shape):
        out = np.empty(shape, np.float32)
        assert self.H.hipMemcpy(out.ctypes.data, self.p, self.n, D2H) == 0
        return out

    def free(self):
        self.H.hipFree(self.p)


def stiff_dev(e, slot, H, u):
    L = semsks.lib()
    L.semsks_stiff_dev.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    ub, fb = DevBuf(H, u.nbytes), DevBuf(H, u.nbytes)
    try:
        ub.put(u)
        semsks._check(L.semsks_stiff_dev(e._h, slot, ub.p, fb.p))
        back = ub.get(u.shape)
        assert np.array_equal(back, u)                # the input is left as it was
        return fb.get(u.shape)
    finally:
        ub.free()
        fb.free()


def load(d):
    old = os.getcwd()
    os.chdir(d)
    try:
        return semsks.Model(["par=run.par"])
    finally:
        os.chdir(old)


def test_no_alloc_stiff_matches_stiff(tmp_path):
    print("case: semsks_stiff_dev, 24x12 tilted + 5x4, two calls", flush=True)
    H = hip()
    big = inputs.small_case(nst=3, interfaces=[([0.0, 50.0], [5.3, 9.1]), None, None],
                            aniso=[None, aniso.hexagonal_tilted(6.5, 3.75, 2.9, 0.05, 30.0, 20.0), None])
    small = inputs.Case(nex=5, eh=2.0, layers_thk=(2.0, 2.0, 4.002), nst=3, nrec=2)
    big.write(os.fspath(tmp_path / "big"))
    small.write(os.fspath(tmp_path / "small"))
    mb, ms = load(tmp_path / "big"), load(tmp_path / "small")
    rng = np.random.default_rng(11)
    with semsks.Engine(max_slots=2) as e:
        sb = e.admit(mb.problem)
        ss = e.admit(ms.problem)
        for slot, m in ((sb, mb), (ss, ms), (sb, mb)):
            u = rng.standard_normal((3, m.nz, m.nx)).astype(np.float32)
            want = e.stiff(slot, u)
            got = stiff_dev(e, slot, H, u)
            assert np.abs(want).max() > 0
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (slot, np.abs(got - want).max())


def test_no_alloc_stiff_refuses_a_mesh_beyond_the_scratch(tmp_path):
    print("case: semsks_stiff_dev, mesh larger than the scratch", flush=True)
    H = hip()
    c = inputs.Case(nex=720, eh=2.0, layers_thk=(20.0, 20.0, 460.001), nst=3, nrec=2)
    c.write(os.fspath(tmp_path))
    m = load(tmp_path)
    assert m.nx * m.nz * 24 > 64 << 20
    u = np.zeros((3, m.nz, m.nx), np.float32)
    with semsks.Engine() as e:
        s = e.admit(m.problem)
        with pytest.raises(semsks.SemsksError) as ei:
            stiff_dev(e, s, H, u)
        assert ei.value.code == 1 and "scratch" in str(ei.value)
