"""The C ABI library loads and exports every function include/*.h declares;
the ctypes layouts match the C structs.  No compute calls (no GPU needed)."""
import ctypes as C
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(mf_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol(pkg):
    lib = pkg._abi.lib()
    declared = _declared()
    assert {"mf_fit_batch", "mf_inplace_fnnls_batch", "mf_fnnls_batch", "mf_linalg_batch",
            "mf_gen_samples"} <= declared
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    bound = {n for n, *_ in pkg._abi.SIGNATURES}
    assert declared <= bound, declared - bound


def test_struct_layouts(pkg):
    a = pkg._abi
    # sizes/offsets of include/mf_abi.h structs (x86-64 SysV)
    assert C.sizeof(a.ChType) == 4 * 14 + 8 * (1 + 100 + 19 + 361)
    assert a.ChType.pederr.offset == 56
    assert C.sizeof(a.Opts) == 56
    assert a.Opts.eps.offset == 16
    assert C.sizeof(a.Batch) == 48 and C.sizeof(a.Result) == 40
    assert a.Batch.pederr.offset == 40
    assert C.sizeof(a.GenSpec) == 8 + 11 * 8
    assert C.sizeof(a.Tune) == 32
    assert pkg._abi.lib().mf_abi_version() == 3


def test_default_opts_are_the_reference_constants(pkg):
    o = pkg._abi.default_opts()
    assert o.eps == 1e-11 and o.max_iter == 1000 and o.max_minimize == 50
    assert o.chi2_tol == 1e-3 and o.precision == 0


def test_product_does_not_link_test_code():
    """The library and the package never load the oracle or the host core."""
    so = (ROOT / "test-regression_amd" / "libmf_amd.so").read_bytes()
    assert b"or_dofit" not in so and b"hc_fit_batch" not in so
    for py in (ROOT / "test-regression_amd").glob("*.py"):
        t = py.read_text()
        for banned in ("oracle_binding", "liboracle", "or_dofit", "hostcore"):
            assert banned not in t, (py, banned)


def test_struct_layouts_match_the_c_compiler(pkg, tmp_path):
    """Every struct of mf_abi.h: size and field offsets as gcc lays them out."""
    import subprocess
    a = pkg._abi
    structs = {"mf_chtype": a.ChType, "mf_opts": a.Opts, "mf_batch": a.Batch,
               "mf_result": a.Result, "mf_timings": a.Timings, "mf_tune": a.Tune}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mf_abi.h"', 'int main(void){']
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, *_ in py._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "lay.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "lay"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), "-o", str(exe), str(src)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                 check=True).stdout.splitlines())
    for cname, py in structs.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for f, *_ in py._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(py, f).offset, (cname, f)
