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
