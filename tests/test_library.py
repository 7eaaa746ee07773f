"""CPU checks of the product library: it builds, loads and exports the C ABI."""
import ctypes
import os
import re

import numpy as np

from radtrancode_amd import _native, synth
from radtrancode_amd.kta import read_kta, write_kta

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "radg.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = "\n".join(l for l in src.splitlines() if not l.lstrip().startswith("#"))
    return sorted(set(re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", src)) -
                  {"if", "sizeof", "while", "for", "return"})


def test_library_loads_and_exports_every_declared_symbol():
    lib = _native.lib()
    names = header_functions()
    assert "cirsradg_wave_" in names and "radg_batch_create" in names
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(_native.EXPORTS) == names


def test_mode_switch_and_error_text_without_a_gpu():
    lib = _native.lib()
    assert lib.radg_set_mode(1) == 0 and lib.radg_get_mode() == 1
    assert lib.radg_set_mode(7) != 0
    assert b"mode" in lib.radg_last_error()
    assert lib.radg_set_mode(0) == 0


def test_kta_round_trip(tmp_path):
    tabs = synth.make_ktables([11], 600.0, 0.8, 5, ng=6, np_=4, nt=3)
    for recl in (4, 8):
        p = str(tmp_path / f"t{recl}.kta")
        write_kta(p, tabs[0], recl=recl)
        back = read_kta(p, recl=recl)
        rtol = 1e-7 if recl == 4 else 0
        assert back.idgas == 11 and back.npoint == 5 and back.ng == 6
        np.testing.assert_allclose(back.k, tabs[0].k, rtol=rtol)
        np.testing.assert_allclose(back.press, tabs[0].press, rtol=rtol)
        np.testing.assert_allclose(back.delg, tabs[0].delg, rtol=rtol)


def test_synthetic_tables_are_monotone_in_g_and_seeded():
    a = synth.make_ktables([11, 27], 600.0, 1.0, 3, ng=20)
    b = synth.make_ktables([11, 27], 600.0, 1.0, 3, ng=20)
    assert np.array_equal(a[0].k, b[0].k)
    assert np.all(np.diff(a[1].k, axis=-1) >= 0)
    assert abs(a[0].delg.sum() - 1) < 1e-14
    press, temp, *_ = synth.make_atmosphere()
    assert press.shape == (120,) and np.all(np.diff(press) < 0)
    assert 40 < temp.min() < temp.max() < 200
