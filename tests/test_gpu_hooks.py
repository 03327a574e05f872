"""The hooks are the engine's extension point: a build of the engine with changed hooks
(tests/hooks/variant_hooks.hpp compiled into multigraph_amd/libmgamd_hookvar.so, as a
maintainer would edit code/user_function.h) must change the results of every path
that runs the changed algorithms, as tests/hooks/check_variant.py predicts.  It runs
in its own process, which loads the variant library instead of libmgamd.so."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
VARIANT = os.path.join(ROOT, "multigraph_amd", "libmgamd_hookvar.so")


@pytest.mark.gpu
def test_changed_hooks_change_every_path():
    assert os.path.exists(VARIANT), "build the variant with __graft_entry__.build()"
    env = dict(os.environ, MG_LIBRARY=VARIANT)
    print("hooks: running tests/hooks/check_variant.py on libmgamd_hookvar.so", flush=True)
    p = subprocess.run([sys.executable, "-u", os.path.join(ROOT, "tests", "hooks", "check_variant.py")], env=env,
                       capture_output=True, text=True, timeout=240)
    print(p.stdout[-4000:], p.stderr[-4000:])
    assert p.returncode == 0


def test_variant_library_exports_the_abi():
    """CPU: the variant build exports the same entry points as libmgamd.so"""
    import ctypes as C

    from test_abi import declared_functions
    L = C.CDLL(VARIANT)
    assert all(hasattr(L, n) for n in declared_functions())
