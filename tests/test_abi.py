"""The C-ABI library loads and exports every symbol include/semsks.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import os
import re

import semsks


def declared_symbols(root):
    src = open(os.path.join(root, "include", "semsks.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(semsks_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol(root):
    L = ctypes.CDLL(semsks.LIB_PATH)
    names = declared_symbols(root)
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(semsks.EXPORTS)


def test_cli_built():
    assert os.access(semsks.CLI_PATH, os.X_OK)


def test_missing_library_fails_loudly(tmp_path):
    """Without the built library the binding raises; there is no fallback."""
    import subprocess
    import sys
    code = ("import semsks, sys\n"
            "semsks.LIB_PATH = %r\n"
            "try:\n    semsks.lib()\nexcept RuntimeError as e:\n    print('raised:', e); sys.exit(0)\n"
            "sys.exit(1)\n") % os.fspath(tmp_path / "libsemsks.so")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, PYTHONPATH=os.pathsep.join(sys.path)))
    assert r.returncode == 0 and "not built" in r.stdout, (r.stdout, r.stderr)


def test_footprint_of_the_plain_update(tmp_path, monkeypatch):
    """semsks_footprint (host code, no GPU): the plain FP32 update
    (SEMSKS_F_NO_KAHAN) keeps no compensation-word planes, 4 bytes per node
    and level of the padded plane fewer than the compensated update."""
    from semsks import inputs
    inputs.small_case().write(os.fspath(tmp_path))
    monkeypatch.chdir(tmp_path)
    m = semsks.Model(["par=run.par"])          # owns the problem's arrays
    p = m.problem
    assert p.flags == 0
    full = semsks.footprint(p)
    p.flags = semsks.F_NO_KAHAN
    plain = semsks.footprint(p)
    # padded plane: column blocks of bw node columns, rounded up to 64 nodes
    nx, nz = p.nex * p.hpn + 1, p.nez * p.hpn + 1
    bw = (64 // (p.hpn + 1) - 1) * p.hpn
    nodes = ((nx + bw - 1) // bw * bw * nz + 63) // 64 * 64
    assert full - plain == 3 * 4 * nodes
