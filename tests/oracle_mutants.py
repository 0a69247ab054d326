"""Copies of the oracle with one planted mistake, built on the CPU.

A pin of the oracle is only worth its name if a plausible mistake in the
oracle fails it.  `mutant()` copies the oracle's C sources into a temporary
directory, replaces one exact piece of text in one file (the planted
mistake), builds the copy with the oracle's own Makefile and makes every
binding of `oracle.oracle` call the copy inside a `with` block.  The shipped
oracle is never touched.
"""
import contextlib
import glob
import os
import shutil
import subprocess

from oracle import oracle as O

ORACLE = os.path.dirname(os.path.abspath(O.__file__))


def build(tmpdir, file, old, new):
    """Build the oracle with `old` replaced by `new` in `file` (exactly one
    occurrence of `old` must exist); returns the path of the library."""
    d = os.path.join(str(tmpdir), "oracle_mutant")
    if os.path.isdir(d):
        shutil.rmtree(d)
    os.makedirs(d)
    for f in glob.glob(os.path.join(ORACLE, "*.c")) + glob.glob(os.path.join(ORACLE, "*.h")) + \
            [os.path.join(ORACLE, "Makefile")]:
        shutil.copy(f, d)
    path = os.path.join(d, file)
    src = open(path).read()
    assert src.count(old) == 1, f"mutation site not unique in {file}: {old!r}"
    open(path, "w").write(src.replace(old, new))
    subprocess.run(["make", "-s", "-C", d], check=True, capture_output=True)
    return os.path.join(d, "liboracle.so")


@contextlib.contextmanager
def mutant(tmpdir, file, old, new):
    """Inside the block the oracle bindings run the mutated build."""
    path = build(tmpdir, file, old, new)
    with O.swapped(path):
        yield


def fails(check, tmpdir, file, old, new):
    """True when `check()` raises AssertionError on the mutated oracle."""
    with mutant(tmpdir, file, old, new):
        try:
            check()
        except AssertionError:
            return True
    return False
