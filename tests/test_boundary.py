"""The C-ABI boundary without a GPU: the library loads and exports every entry point the header
declares, the ctypes structures match the C layout, the device code has no contracted FP64
multiply-add, and the NumPy semantics the device helpers restate hold on special values."""

import ctypes
import pathlib
import re
import subprocess

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "tm5hip.h"


def declared():
    return re.findall(r"^int (tm5_\w+)\(", HEADER.read_text(), flags=re.M)


def test_library_exports_every_declared_entry_point():
    from pydyn_amd import _lib

    names = declared()
    assert len(names) >= 20
    lib = _lib.lib()
    for name in names:
        assert hasattr(lib, name), name
        assert name in _lib._SIGNATURES, f"{name} has no ctypes prototype"
    assert set(_lib._SIGNATURES) == set(names)


def test_ctypes_structures_match_the_c_layout(tmp_path):
    from pydyn_amd import _lib

    probe = tmp_path / "probe.c"
    checks = {
        "tm5_view": _lib.View,
        "tm5_tracers": _lib.Tracers,
        "tm5_iview": _lib.IView,
        "tm5_region_desc": _lib.RegionDesc,
        "tm5_hview": _lib.HView,
        "tm5_zregion": _lib.ZRegion,
        "tm5_zphysics": _lib.ZPhysics,
        "tm5_tree_desc": _lib.TreeDesc,
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "tm5hip.h"', "int main(void) {"]
    for cname, cls in checks.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    probe.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(probe), "-o", str(exe)], check=True)
    out = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split("\n") if line)
    for cname, cls in checks.items():
        assert int(out[cname]) == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert int(out[f"{cname}.{fname}"]) == getattr(cls, fname).offset, f"{cname}.{fname}"


@pytest.mark.parametrize("source", ["advect.hip", "convec.hip", "column.hip", "lapack.hip", "physics.hip", "zoom.hip", "tree.hip",
                                    "adjoint.hip", "api.hip"])
def test_no_contracted_fp64_fma_in_device_code(source, tmp_path):
    """Every v_fma_f64 / v_fmac_f64 in the gfx950 code belongs to a correctly rounded division
    (v_div_scale .. v_div_fmas .. v_div_fixup, five FMAs each) or square root (v_rsq_f64 and
    seven FMAs of refinement, fdlibm pow's y = 0.5 case): the kernels' own adds and multiplies
    are never fused, so they round like NumPy's."""
    from pydyn_amd import build as B

    asm = tmp_path / "out.s"
    flags = [f for f in B.FLAGS if f not in ("-shared", "-fPIC")]
    cmd = [B.hipcc(), *flags, f"-I{ROOT / 'include'}", f"-I{B.CSRC}", "--offload-device-only", "-S",
           str(B.CSRC / source), "-o", str(asm)]  # fmt: skip
    subprocess.run(cmd, check=True, capture_output=True)
    text = asm.read_text()
    fma = len(re.findall(r"\bv_fmac?_f64", text))
    divisions = len(re.findall(r"\bv_div_fmas_f64", text))
    roots = len(re.findall(r"\bv_rsq_f64", text))
    assert fma == 5 * divisions + 7 * roots, f"{source}: {fma} FP64 FMAs for {divisions} divisions, {roots} roots"
    assert not re.search(r"\bv_pk_(fma|add|mul)_f32", text), f"{source}: packed FP32 instructions"


def test_numpy_clip_and_maximum_semantics_the_device_restates():
    """np.clip with array bounds is max-by-'>' then min-by-'<', NaN first; np.maximum(0, x)
    keeps x on a tie.  tm5_device.h's limit() and max0() are these formulas."""
    vals = [-0.0, 0.0, 1.0, -1.0, np.nan, np.inf, -np.inf, 5e-324]
    for n in (1, 7, 64, 1000):
        for s in vals:
            for r in vals:
                got = np.clip(np.full(n, s), -np.abs(np.full(n, r)), np.abs(np.full(n, r)))[0]
                hi = abs(r)
                lo = -hi
                t = s if np.isnan(s) else (s if s > lo else lo)
                want = t if np.isnan(t) else (t if t < hi else hi)
                assert np.array_equal(np.array([got]).view(np.int64), np.array([want]).view(np.int64)) or (
                    np.isnan(got) and np.isnan(want)
                ), (n, s, r, got, want)
        x = np.array([-0.0, 0.0, np.nan, -1.0, 2.0] * n)
        got = np.maximum(0.0, x)
        want = np.array([0.0 if 0.0 > v else v for v in x])
        assert np.array_equal(got.view(np.int64), want.view(np.int64)) or np.array_equal(np.isnan(got), np.isnan(want))


def test_meteo_inputs_are_reproducible():
    """The generator's inputs for a seed are fixed bits (the golden runs were made from them)."""
    import hashlib

    from pydyn_amd import meteo

    met = meteo.Meteo(meteo.grid("tiny"), seed=3)
    h = hashlib.sha1()
    for k in ("m", "pu0", "pv0", "entu", "detu", "entd", "detd", "kvh", "gph", "emis", "loss", "rm"):
        h.update(np.ascontiguousarray(getattr(met, k)).tobytes())
    h.update(met.cloud_top.tobytes())
    h.update(met.cloud_lfs.tobytes())
    assert h.hexdigest() == (ROOT / "tests" / "golden" / "meteo_tiny_sha1.txt").read_text().strip()
