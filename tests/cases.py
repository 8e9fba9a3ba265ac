"""Shared synthetic cases for the tests (written in the reference's formats)."""
import os

from semsks import inputs


def write_case(tmpdir, **kw):
    c = inputs.small_case(**kw)
    d = os.fspath(tmpdir)
    D = c.write(d)
    return c, D, d
