"""The C-ABI library loads on a host without a GPU and exports every
entry point include/arwen_dycore.h declares (no device calls here), and
the ctypes struct layouts match the header's field order and sizes."""
import ctypes
import re
from pathlib import Path

import pytest

from arwen_amd import _lib

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "arwen_dycore.h"


def _declared():
    text = HEADER.read_text()
    return sorted(set(re.findall(
        r"^(?:int32_t|const char \*|void|double)\s*\*?\s*(arw(?:64)?_\w+)\(",
        text, flags=re.M)))


def _lib_or_skip():
    if not _lib.LIB_PATH.exists():
        pytest.skip("library not built (run __graft_entry__.build())")
    return ctypes.CDLL(str(_lib.LIB_PATH))


def test_header_declares_entry_points():
    names = _declared()
    assert "arw_step" in names and len(names) >= 20


def test_library_exports_every_declared_symbol():
    lib = _lib_or_skip()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_the_header():
    assert sorted(_lib.exported_symbols()) == _declared()


def test_abi_version_matches_binding():
    lib = _lib_or_skip()
    lib.arw_abi_version.restype = ctypes.c_int32
    text = HEADER.read_text()
    ver = int(re.search(r"#define ARW_ABI_VERSION (\d+)", text).group(1))
    assert lib.arw_abi_version() == ver == _lib.ABI_VERSION


def test_species_table_size_matches_header_and_schemes():
    """The domain's extra-species pointer tables hold every scheme's set."""
    from arwen_amd.config import EXTRA_SPECIES, MAX_EXTRA_SPECIES
    text = HEADER.read_text()
    n = int(re.search(r"#define ARW_MAX_EXTRA (\d+)", text).group(1))
    assert n == _lib.MAX_EXTRA == MAX_EXTRA_SPECIES
    assert max(len(v) for v in EXTRA_SPECIES.values()) <= n


def test_adjoint_refuses_mixed_phase_configs():
    from arwen_amd import adjoint
    from arwen_amd.state import DomainState
    from arwen_amd import cases
    case = cases.mixed_phase(6, nx=8, ny=6, nz=5)
    st = DomainState.from_host(case, device="cpu")
    with pytest.raises(_lib.ArwenError, match="mp_physics"):
        adjoint._check_adjoint_cfg(st, case.cfg)


def _struct_fields(name):
    text = HEADER.read_text()
    body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name),
                     text, flags=re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        m = re.match(r"(?:const\s+)?(int32_t|uint32_t|float|double)\s+(.*)",
                     decl, flags=re.S)
        base, rest = m.group(1), m.group(2)
        for part in rest.split(","):
            part = part.strip()
            ptr = part.startswith("*")
            name = part.lstrip("*").strip()
            arr = re.match(r"(\w+)\[(\w+)\]", name)
            if arr:                       # pointer table: name[ARW_MAX_EXTRA]
                fields.append((arr.group(1), "ptr[]"))
                continue
            fields.append((name, "ptr" if ptr else base))
    return fields


@pytest.mark.parametrize("cname,pyt", [("arw_domain", _lib.ArwDomain),
                                       ("arw64_domain", _lib.ArwDomain64),
                                       ("arw_params", _lib.ArwParams),
                                       ("arw_stage_adj", _lib.ArwStageAdj),
                                       ("arw64_stage_adj", _lib.ArwStageAdj64),
                                       ("arw_init_adj", _lib.ArwInitAdj),
                                       ("arw64_init_adj", _lib.ArwInitAdj64),
                                       ("arw_finish_adj", _lib.ArwFinishAdj),
                                       ("arw64_finish_adj", _lib.ArwFinishAdj64)])
def test_ctypes_layout_matches_header(cname, pyt):
    kinds = {"ptr": ctypes.c_void_p, "int32_t": ctypes.c_int32,
             "uint32_t": ctypes.c_uint32,
             "float": ctypes.c_float, "double": ctypes.c_double,
             "ptr[]": ctypes.c_void_p * _lib.MAX_EXTRA}
    want = [(n, kinds[k]) for n, k in _struct_fields(cname)]
    have = [(n, t) for n, t in pyt._fields_]
    assert [n for n, _ in have] == [n for n, _ in want]
    assert all(ctypes.sizeof(a) == ctypes.sizeof(b)
               for (_, a), (_, b) in zip(have, want))


def _c_layout(struct, fields, tmp_path):
    """sizeof and offsetof from the C compiler for the header's struct."""
    import subprocess
    src = tmp_path / "layout.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>',
             '#include "arwen_dycore.h"', 'int main(void) {',
             f'  printf("%zu\\n", sizeof({struct}));']
    lines += [f'  printf("%zu\\n", offsetof({struct}, {f}));' for f in fields]
    lines += ['  return 0;', '}']
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", str(ROOT / "include"), str(src),
                    "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True,
                         text=True).stdout.split()
    return [int(v) for v in out]


@pytest.mark.parametrize("cname,pyt", [
    ("arw_tracer_ckpt", _lib.ArwTracerCkpt),
    ("arw64_tracer_ckpt", _lib.ArwTracerCkpt64),
    ("arw_acoustic_adj", _lib.ArwAcousticAdj),
    ("arw64_acoustic_adj", _lib.ArwAcousticAdj64),
    ("arw_domain", _lib.ArwDomain), ("arw64_domain", _lib.ArwDomain64),
    ("arw_params", _lib.ArwParams),
    ("arw_stage_adj", _lib.ArwStageAdj), ("arw_init_adj", _lib.ArwInitAdj),
    ("arw_acoustic_ref", _lib.ArwAcousticRef),
    ("arw64_acoustic_ref", _lib.ArwAcousticRef64),
    ("arw_acoustic_fwd", _lib.ArwAcousticFwd),
    ("arw64_acoustic_fwd", _lib.ArwAcousticFwd64),
    ("arw_health_field", _lib.ArwHealthField)])
def test_ctypes_offsets_match_the_c_compiler(cname, pyt, tmp_path):
    names = [n for n, _ in pyt._fields_]
    got = _c_layout(cname, names, tmp_path)
    assert got[0] == ctypes.sizeof(pyt)
    assert got[1:] == [getattr(pyt, n).offset for n in names]


def test_held_mixing_scratch_and_params():
    """km_opt = 4 and the filter on a moist state: the domain view carries
    the Smagorinsky work area (9 nz ny nx + 5 (nz+1) ny nx) and one held
    slot per species; the filter's species mask follows moist_mix6_off."""
    from arwen_amd import cases, dycore
    from arwen_amd.state import DomainState
    case = cases.mixed_phase(10, nx=8, ny=6, nz=5)
    cfg = case.cfg.with_(km_opt=4, diff_6th_opt=2)
    st = DomainState.from_host(case, device="cpu")
    d = st.domain_struct(diff6=True, smag=True)
    nz, ny, nx = 5, 6, 8
    n, nw = nz * ny * nx, (nz + 1) * ny * nx
    assert st._scratch["arw_sm_buf"].numel() == 9 * n + 5 * nw
    nq = 3 + len(st.extra_species)
    assert tuple(st._scratch["arw_held_q"].shape) == (nq, nz, ny, nx)
    assert d.sm_buf and d.held_q and d.d6_u and d.d6_t
    p = dycore.params(cfg)
    assert p.km_opt == 4 and p.isfflx == 1 and abs(p.c_s - 0.25) < 1e-15
    assert p.diff6_species == (1 << nq) - 1
    off = dycore.params(cfg.with_(moist_mix6_off=True)).diff6_species
    names = ("qv", "qc", "qr") + st.extra_species
    for e, name in enumerate(names):
        kept = bool((off >> e) & 1)
        assert kept == (name not in dycore.MOIST_ARRAY), name
    assert dycore.params(case.cfg).diff6_species == 0
