"""Every name a function of the committed Python reads must be bound
somewhere it can see: in the function, an enclosing function, the module
or builtins.  A name that is bound nowhere raises NameError only when that
line runs (a bench workload that is rarely run, an error path), so this
walks the compiler's symbol tables of each file instead of running it."""
import builtins
import os
import symtable

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _sources():
    out = [os.path.join(ROOT, f) for f in ("bench.py", "__graft_entry__.py")]
    for sub in ("efdcplus8.5_amd/efdc_amd", "oracle", "tests", "tests/golden"):
        d = os.path.join(ROOT, sub)
        if os.path.isdir(d):
            out += [os.path.join(d, f) for f in sorted(os.listdir(d)) if f.endswith(".py")]
    return out


def unbound_names(path):
    """(scope, name) pairs read in a function or class of `path` that no
    scope binds."""
    src = open(path).read()
    top = symtable.symtable(src, path, "exec")
    module_names = {s.get_name() for s in top.get_symbols()
                    if s.is_assigned() or s.is_imported() or s.is_namespace()}
    star = any(s.get_name() == "*" for s in top.get_symbols())
    known = module_names | set(dir(builtins)) | {"__file__", "__name__", "__doc__"}
    bad = []

    def walk(t):
        for s in t.get_symbols():
            if not s.is_referenced():
                continue
            if s.is_global() and not s.is_declared_global() or \
                    (t.get_type() == "module" and not (s.is_assigned() or s.is_imported())):
                if s.get_name() not in known and not star:
                    bad.append((t.get_name(), s.get_name()))
        for c in t.get_children():
            walk(c)

    walk(top)
    return bad


@pytest.mark.parametrize("path", _sources(), ids=lambda p: os.path.relpath(p, ROOT))
def test_no_unbound_names(path):
    assert unbound_names(path) == []


def test_the_check_sees_an_unbound_name(tmp_path):
    f = tmp_path / "m.py"
    f.write_text("import os\n\ndef run(args):\n    n = args.n\n    return f'x{nx}' + os.sep\n")
    assert unbound_names(str(f)) == [("run", "nx")]
