"""Every name a host-side Python function reads resolves: a local, an
enclosing function's name, a module-level binding or a builtin.  Catches a
function that refers to a variable of another function (a NameError that
only a GPU run of that code path would show otherwise)."""
import builtins
import os
import symtable

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FILES = ["bench.py", "__graft_entry__.py", "oracle/__init__.py",
         "nfft.jl_amd/nfftm/__init__.py", "nfft.jl_amd/nfftm/_lib.py", "nfft.jl_amd/nfftm/tools.py",
         "nfft.jl_amd/nfftm/autograd.py"]


def unresolved(path):
    top = symtable.symtable(open(path).read(), path, "exec")
    module = {s.get_name() for s in top.get_symbols() if s.is_assigned() or s.is_imported()
              or s.is_namespace()}
    known = module | set(dir(builtins)) | {"__file__", "__name__"}
    bad = []

    def walk(t):
        for c in t.get_children():
            if c.get_type() == "function":
                for s in c.get_symbols():
                    if s.is_referenced() and s.is_global() and not s.is_declared_global() \
                            and s.get_name() not in known:
                        bad.append(f"{c.get_name()}: {s.get_name()}")
            walk(c)

    walk(top)
    return bad


@pytest.mark.parametrize("rel", FILES)
def test_names_resolve(rel):
    path = os.path.join(REPO, rel)
    if not os.path.exists(path):
        pytest.skip(f"{rel} absent")
    assert unresolved(path) == []
