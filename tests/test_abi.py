"""CPU checks of the C-ABI library: it loads, exports every entry point
include/thindb_gpu.h declares, and the struct layouts the bindings use match
the header.  No compute call is made (no GPU here)."""
import ctypes as C
import re
import subprocess
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
    assert lib.tdbg_abi_version() == gpu.ABI_VERSION == 3


STRUCTS = {  # header type -> ctypes binding
    "tdbg_column_key": gpu.ColumnKey, "tdbg_block_stats": gpu.BlockStats, "tdbg_rg_ref": gpu.RgRef,
    "tdbg_key_part": gpu.KeyPart, "tdbg_agg_spec": gpu.AggSpecC, "tdbg_predicate": gpu.PredicateC,
    "tdbg_order_spec": gpu.OrderSpecC, "tdbg_having_node": gpu.HavingNodeC, "tdbg_filter_node": gpu.FilterNodeC,
    "tdbg_plan_desc": gpu.PlanDesc, "tdbg_column_view": gpu.ColumnView,
    "tdbg_exec_stats": gpu.ExecStats, "tdbg_batch_stats": gpu.BatchStats, "tdbg_batch_route": gpu.BatchRoute,
    "tdbg_server_config": gpu.ServerConfig, "tdbg_ticket_report": gpu.TicketReport,
    "tdbg_server_counters": gpu.ServerCounters,
}


def test_struct_layouts(tmp_path):
    """Every struct's size and field offsets as the C compiler lays out the
    header equal the ctypes bindings'."""
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for cname, cls in STRUCTS.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", str(src), "-o", str(exe)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                 text=True).stdout.splitlines())
    for cname, cls in STRUCTS.items():
        assert int(got[cname]) == C.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(cls, f).offset, (cname, f)


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
