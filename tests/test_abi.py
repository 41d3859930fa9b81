"""CPU checks of the C-ABI library: it loads, exports every entry point
include/thindb_gpu.h declares, and the struct layouts the bindings use match
the header.  No compute call is made (no GPU here)."""
import ctypes as C
import re
from pathlib import Path

from thindb_amd import gpu

HEADER = Path(__file__).resolve().parent.parent / "include" / "thindb_gpu.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tdbg_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported_and_bound():
    lib = C.CDLL(str(gpu._LIB_PATH))
    names = declared_functions()
    assert len(names) >= 18
    for n in names:
        assert hasattr(lib, n), n
        assert n in gpu.EXPORTS, f"{n} has no ctypes binding"


def test_abi_version_handshake():
    lib = gpu.load()
    assert lib.tdbg_abi_version() == gpu.ABI_VERSION == 1


def test_struct_layouts():
    # sizes follow from the header's field order (natural alignment)
    assert C.sizeof(gpu.ColumnKey) == 24
    assert C.sizeof(gpu.BlockStats) == 40
    assert C.sizeof(gpu.RgRef) == 16
    assert C.sizeof(gpu.KeyPart) == 8
    assert C.sizeof(gpu.AggSpecC) == 16
    assert C.sizeof(gpu.PredicateC) == 32
    assert C.sizeof(gpu.OrderSpecC) == 8
    assert C.sizeof(gpu.PlanDesc) == 80
    assert C.sizeof(gpu.ColumnView) == 40
    assert C.sizeof(gpu.ExecStats) == 56


def test_enums_match_header():
    text = HEADER.read_text()
    for e in gpu.Type:
        assert re.search(rf"TDBG_{e.name}\s*=\s*{int(e)}\b", text), e
    for e in gpu.PredOp:
        assert re.search(rf"TDBG_{e.name}\s*=\s*{int(e)}\b", text), e
    for e in gpu.AggOp:
        assert re.search(rf"TDBG_AGG_{e.name}\s*=\s*{int(e)}\b", text), e
    for e in gpu.Status:
        name = {"OK": "TDBG_OK", "PENDING": "TDBG_PENDING"}.get(e.name, f"TDBG_ERR_{e.name}")
        assert re.search(rf"{name}\s*=\s*{int(e)}\b", text), e
