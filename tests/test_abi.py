"""The C ABI: libsphdem.so loads and exports every function of include/sphdem.h
(no device work; runs without a GPU), and the ctypes structures match the
header's field lists."""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

from warp_sph_gpt_amd import _lib

ROOT = Path(__file__).resolve().parents[1]
HEADER = (ROOT / "include" / "sphdem.h").read_text()


def declared_functions():
    return sorted(set(re.findall(r"^(?:int|long|const char \*)\s*(sphdem_\w+)\s*\(", HEADER,
                                 re.M)))


def struct_fields(name):
    end = re.search(r"\}\s*" + name + r"\s*;", HEADER).start()
    start = HEADER.rindex("typedef struct {", 0, end) + len("typedef struct {")
    body = re.sub(r"/\*.*?\*/", "", HEADER[start:end], flags=re.S)
    out = []
    for decl in body.split(";"):
        words = [w for w in re.split(r"[\s,*]+", decl) if w]
        out.extend(w.split("[")[0] for w in words if w not in ("int", "double", "void", "const"))
    return out


def test_library_exports_every_declared_function():
    if not _lib.LIB_PATH.exists():
        pytest.skip("libsphdem.so is not built (run python -m warp_sph_gpt_amd.build)")
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(_lib.exported_symbols()), set(names) ^ set(_lib.exported_symbols())
    assert lib.sphdem_version() == 1
    assert lib.sphdem_scan_scratch(4096) >= 2


@pytest.mark.parametrize("cname,pyname", [
    ("sphdem_params_t", "Params"), ("sphdem_sph_t", "SphView"), ("sphdem_bnd_t", "BndView"),
    ("sphdem_dem_t", "DemView"), ("sphdem_dembnd_t", "DemBndView"),
    ("sphdem_grid_t", "GridStruct"), ("sphdem_work_t", "WorkStruct"),
    ("sphdem_tuning_t", "TuningStruct")])
def test_ctypes_structures_follow_the_header(cname, pyname):
    fields = struct_fields(cname)
    got = [f[0] for f in getattr(_lib, pyname)._fields_]
    assert got == fields, (cname, fields, got)


# the repository's integration notes, whose first python block is the
# reference-side ctypes binding a maintainer would add
INTEGRATION = (ROOT / "INTEGRATION.md").read_text()


def integration_struct(name):
    """Run the ctypes class `name` of the integration notes' reference-side
    binding (its definition only) and return it."""
    code = re.search(r"```python\n(.*?)```", INTEGRATION, re.S).group(1)
    start = code.index(f"class {name}(ctypes.Structure)")
    end = code.index("\n\n", start)
    ns = {"ctypes": ctypes}
    exec(code[start:end], ns)
    return ns[name]


def test_integration_grid_binding_matches_the_header():
    """The documented ctypes Grid has every field of sphdem_grid_t, in order,
    with the library's own binding's layout (size and offsets)."""
    Grid = integration_struct("Grid")
    assert [f[0] for f in Grid._fields_] == struct_fields("sphdem_grid_t")
    assert ctypes.sizeof(Grid) == ctypes.sizeof(_lib.GridStruct)
    for (name, _), (name2, _) in zip(Grid._fields_, _lib.GridStruct._fields_):
        assert getattr(Grid, name).offset == getattr(_lib.GridStruct, name2).offset, name


def test_tuning_round_trips_and_rejects_out_of_range_values():
    """Launch shapes are host-side state: set, read back, and invalid
    values refused with the library's error (no device work)."""
    saved = _lib.get_tuning()
    try:
        _lib.set_tuning({"por_parts": 7, "por_x_tile": 3, "dem_beside": 0,
                         "block": {"porosity": 64}, "por_queue": 0, "wall_beside": 0,
                         "force_beside": 0})
        t = _lib.get_tuning()
        assert (t["por_parts"], t["por_x_tile"], t["dem_beside"], t["por_queue"],
                t["wall_beside"], t["force_beside"]) == (7, 3, 0, 0, 0, 0)
        assert t["block"]["porosity"] == 64
        for bad in ({"por_parts": 0}, {"por_parts": 17}, {"por_x_tile": 7},
                    {"dem_beside": 2}, {"block": {"porosity": 96}}, {"por_queue": 2}, {"wall_beside": 2},
                    {"force_beside": 2},
                    {"por_queue": -1}):
            with pytest.raises(ValueError, match="sphdem_tuning_set"):
                _lib.set_tuning(bad)
        assert _lib.get_tuning() == t
    finally:
        _lib.set_tuning(saved)
