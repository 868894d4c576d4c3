"""CPU checks of the C-ABI library: it loads, exports every symbol that
include/oocdag.h declares, and its host-side logic (panel planner, storage
size) agrees with the oracle.  No GPU compute is called here."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "oocdag.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(oocdag_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol(ooc):
    L = ooc.lib()
    names = declared_symbols()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name
    # and the Python binding knows the signature of each one
    assert set(names) == set(ooc.SIGNATURES)


@pytest.mark.parametrize("n,nb,mem", [
    (4096, 256, 28 << 20), (4096, 256, 32 << 20), (196608, 1024, 64 * 10**9),
    (196608, 1024, 64 << 30), (196608, 1024, 160 << 30), (1000, 64, 3 << 20), (777, 128, 2 << 20),
])
def test_single_x_plan_is_reference_find_Yn(ooc, ora, n, nb, mem):
    assert ooc.panel_plan(n, nb, mem, nx=1) == ora.panel_plan(n, nb, mem)


@pytest.mark.parametrize("n,nb,mem", [
    (4096, 256, 28 << 20), (4096, 256, 32 << 20), (196608, 1024, 64 * 10**9), (1000, 64, 3 << 20),
    (4096, 256, 16 << 20),
])
def test_double_buffered_plan_fits_budget(ooc, n, nb, mem):
    bb = -(-n // nb)
    memBlock = mem // 8 // nb // nb
    widths = ooc.panel_plan(n, nb, mem)
    assert widths is not None and sum(widths) == bb
    nx = 2 if ooc.panel_plan(n, nb, mem, nx=2) == widths else 1
    jb = 0
    for y in widths:
        ym = bb - jb
        trap = y * ym - y * (y - 1) // 2
        assert trap + (nx * ym if jb > 0 else 0) <= memBlock
        jb += y


def test_budget_below_two_columns_is_an_error(ooc):
    assert ooc.panel_plan(4096, 256, 15 << 19) is None  # w = memBlock/bb - 1 = 0


def test_tiles_elems(ooc):
    for n, nb in [(4096, 256), (1000, 64), (196608, 1024)]:
        bb = -(-n // nb)
        assert ooc.tiles_elems(n, nb) == bb * (bb + 1) // 2 * nb * nb


def test_no_device_is_reported_not_faked(ooc):
    if ooc.device_count() > 0:
        pytest.skip("a GPU is visible")
    A = np.eye(128, order="F")
    with pytest.raises(ooc.OocdagError):
        ooc.cholesky(A, 64, 1 << 30)


def test_product_never_touches_the_oracle():
    pkg = os.path.join(ROOT, "oocdag_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in text and "liboracle" not in text, f
                assert "ora_" not in text, f
