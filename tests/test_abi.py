"""The C-ABI library loads and exports every entry point include/*.h declares."""
import os
import re

import pytest

from enscdynamics_amd import engine as eng

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(ensc_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_the_api():
    names = declared_functions()
    assert "ensc_engine_create" in names and "ensc_engine_run" in names
    assert set(eng.EXPORTS) <= set(names)


def test_library_exports_every_declared_symbol():
    lib = eng.load_library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_create_without_gpu_fails_loudly():
    if eng.device_count() > 0:
        pytest.skip("a GPU is visible")
    from enscdynamics_amd import decks
    m = decks.config1()
    with pytest.raises(eng.EnscError):
        eng.Engine(m)
