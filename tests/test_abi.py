"""The C-ABI library loads and exports every entry point include/nbody_gpu.h declares (CPU only:
no compute call is made here)."""
import ctypes as C
import os
import re

import pytest

from nbodygo_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))
           if f.endswith(".h")]


def declared():
    names = []
    for h in HEADERS:
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(nb[a-z_0-9]*)\s*\(", src, re.M)
    return sorted(set(names))


def test_every_declared_symbol_is_exported():
    names = declared()
    assert len(names) >= 15
    lib = C.CDLL(L.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_the_header():
    bound = {s[0] for s in L.SIGNATURES}
    assert set(declared()) <= bound | set(getattr(L, "EXTRA_SIGNATURES_NAMES", set()))


def test_struct_layouts_match_header(tmp_path):
    """sizeof/offsetof of every struct field as gcc lays out the header == the ctypes mirror."""
    import subprocess
    structs = {"nb_params": L.NbParams, "nb_stats": L.NbStats, "nb_diag": L.NbDiag,
               "nb_sur_params": L.NbSurParams, "nb_sur_stats": L.NbSurStats,
               "nb_body": L.NbBody, "nb_bh_params": L.NbBhParams, "nb_bh_stats": L.NbBhStats,
               "nb_event_stats": L.NbEventStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "nbody_gpu.h"',
             "int main(void){"]
    for cname, cls in structs.items():
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for f in cls._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f[0], cname,
                                                                          f[0]))
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)])
    got = dict(l.split() for l in subprocess.check_output([str(exe)]).decode().splitlines())
    for cname, cls in structs.items():
        assert int(got[cname]) == C.sizeof(cls), cname
        for f in cls._fields_:
            assert int(got["%s.%s" % (cname, f[0])]) == getattr(cls, f[0]).offset, (cname, f[0])


def test_params_defaults_are_reference_semantics():
    p = L.NbParams()
    assert L.lib.nb_params_init(C.byref(p)) == L.NB_OK
    assert p.G == 6.673e-11 and p.eps == 0.0
    assert p.overlap_cutoff == 1 and p.collisions == 1 and p.precision == L.NB_PREC_FP32
    assert L.lib.nb_params_init(None) == L.NB_E_INVALID


def test_null_arguments_are_rejected_without_a_device():
    assert L.lib.nb_create(None, None) == L.NB_E_INVALID
    out = C.c_void_p()
    assert L.lib.nb_create(None, C.byref(out)) == L.NB_E_INVALID
    assert not out.value
    assert b"NULL" in L.lib.nb_last_error(None)
    p = L.NbParams()
    L.lib.nb_params_init(C.byref(p))
    p.eps = -1.0
    assert L.lib.nb_create(C.byref(p), C.byref(out)) == L.NB_E_INVALID
    p.eps = 1e-20  # eps^2 underflows FP32
    assert L.lib.nb_create(C.byref(p), C.byref(out)) == L.NB_E_INVALID
    assert b"underflows" in L.lib.nb_last_error(None)
    for fn in ("nb_step", "nb_set_timing"):
        assert getattr(L.lib, fn) is not None
    assert L.lib.nb_set_timing(None, 1) == L.NB_E_INVALID
    assert L.lib.nb_destroy(None) is None


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(ImportError):
        L.load_library(str(tmp_path / "libnbgpu.so"))


def test_surrogate_params_defaults_and_null_handling():
    p = L.NbSurParams()
    assert L.lib.nb_sur_params_init(C.byref(p)) == L.NB_OK
    assert (p.grid, p.shell, p.r0, p.use_mlp) == (32, 2, 0.0, 1)
    assert (p.check_every, p.check_samples, p.tol) == (1, 256, 1e-3)
    assert L.lib.nb_sur_params_init(None) == L.NB_E_INVALID
    assert L.lib.nb_sur_enable(None, C.byref(p)) == L.NB_E_INVALID
    assert L.lib.nb_sur_disable(None) == L.NB_E_INVALID
    assert L.lib.nb_sur_collect(None, 10, 0) == L.NB_E_INVALID
    assert L.lib.nb_sur_fit(None, 1, 64, 1e-3, None) == L.NB_E_INVALID
    st = L.NbSurStats()
    assert L.lib.nb_sur_get_stats(None, C.byref(st)) == L.NB_E_INVALID
    assert L.lib.nb_sur_save(None, b"x") == L.NB_E_INVALID
    assert L.lib.nb_sur_load(None, b"x") == L.NB_E_INVALID


def test_lifecycle_null_handling():
    b = L.NbBody()
    n = C.c_int64()
    assert L.lib.nb_add_bodies(None, 0, *([None] * 12)) == L.NB_E_INVALID
    assert L.lib.nb_cycle(None, 1.0, C.byref(n)) == L.NB_E_INVALID
    assert L.lib.nb_remove_bodies(None, 1, C.byref(n)) == L.NB_E_INVALID
    assert L.lib.nb_get_body(None, 0, C.byref(b)) == L.NB_E_INVALID
    assert L.lib.nb_set_body(None, 0, C.byref(b)) == L.NB_E_INVALID
    assert L.lib.nb_store_bodies(None, *([None] * 13)) == L.NB_E_INVALID
    assert L.lib.nb_set_frag_steps(None, None) == L.NB_E_INVALID
    ev = L.NbEventStats()
    assert L.lib.nb_get_event_stats(None, C.byref(ev)) == L.NB_E_INVALID


def test_barnes_hut_params_defaults_and_null_handling():
    p = L.NbBhParams()
    assert L.lib.nb_bh_params_init(C.byref(p)) == L.NB_OK
    assert (p.theta, p.softening, p.max_depth) == (0.3, 1e-4, 21)  # octree.go:9, :14
    assert L.lib.nb_bh_params_init(None) == L.NB_E_INVALID
    assert L.lib.nb_bh_enable(None, C.byref(p)) == L.NB_E_INVALID
    assert L.lib.nb_bh_disable(None) == L.NB_E_INVALID
    st = L.NbBhStats()
    assert L.lib.nb_bh_get_stats(None, C.byref(st)) == L.NB_E_INVALID
    assert L.lib.nb_bh_get_depths(None, None) == L.NB_E_INVALID
