"""The C-ABI library loads on a CPU-only host and exports every symbol include/seqvp.h declares."""
import re
from pathlib import Path

import pytest

from seq_vp_3d_amd import _lib

HEADER = Path(__file__).resolve().parents[1] / "include" / "seqvp.h"


def declared():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^(?:int|size_t|const char\*)\s+(seqvp_\w+)\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert len(names) >= 20
    assert "seqvp_q1_soft_energy" in names and "seqvp_dst_poisson" in names


@pytest.mark.skipif(not _lib.LIB_PATH.exists(), reason="library not built")
def test_every_declared_symbol_is_exported_and_bound():
    lib = _lib.load()
    for name in declared():
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, name


@pytest.mark.skipif(not _lib.LIB_PATH.exists(), reason="library not built")
def test_host_only_queries():
    lib = _lib.load()
    assert lib.seqvp_version().startswith(b"seqvp-hip gfx950")
    nz, ny, nx = 160, 192, 224
    ncell = (nz - 1) * (ny - 1) * (nx - 1)
    assert _lib.workspace_bytes(2, (nz, ny, nx)) == 24 * 8 * ncell
    # the sine matrices are stored folded in 16-output strips: ceil(n/2) rows rounded up to 16, 16 columns per strip
    # (ceil(ceil(n/2)/16) even strips and ceil(floor(n/2)/16) odd ones)
    assert _lib.workspace_bytes(6, (nz, ny, nx)) == 8 * (3 * 158 * 190 * 222 + 112 * 224 + 96 * 192 + 80 * 160)
    # the T66 sweep: partials of its cubic pass (at most 512 workgroups), the listed-wave count (256 bytes), one
    # list slot per wave of 64 cells
    assert _lib.workspace_bytes(5, (nz, ny, nx)) == 512 * 3 * 8 + 256 + 8 * ((ncell + 63) // 64)
    assert _lib.workspace_bytes(0, (1, 5, 5)) == 0
