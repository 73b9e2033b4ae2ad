"""The C-ABI library loads and exports every entry point declared in
include/igm_amd.h; argument checks fail cleanly without touching a GPU."""
import ctypes
import os
import re

import pytest

from igm_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "igm_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(igm_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = L.load()
    names = declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(L.declared_symbols()) == names  # the ctypes table covers the header


def test_argument_errors_are_reported():
    lib = L.load()
    rc = lib.igm_energy_coeffs_f32(None, None, None, None, None, 0, 0, 0, None, 1, 4, 4, None, None, None, None)
    assert rc < 0
    assert b"null pointer" in lib.igm_last_error()
    with pytest.raises(L.IgmError):
        L.call("igm_energy_finalize", None, 1, 4, 4, None, None, None, 0, None)


def test_workspace_sizes():
    lib = L.load()
    # 1000 x 1000: the one-iteration passes split 17 strips of 62 columns into
    # 120 row blocks (2040 waves, 510 blocks of 4, 3 partial sums each); the
    # two-iteration pass 17 strips of 60 into 60 (1020 waves, 255 blocks, 6
    # each): 1530 doubles for either
    assert lib.igm_energy_workspace_doubles(1, 1000, 1000) == 1530
    # a batch of 1000: one row block per strip, 5 blocks per entry (more than
    # the two-iteration pass's 1530)
    assert lib.igm_energy_workspace_doubles(1000, 1000, 1000) == 3 * 1000 * 5
    assert lib.igm_energy_workspace_doubles(0, 10, 10) == 0
    assert lib.igm_wgrad_workspace_floats(9, 10) == 10 * (9 * 1024 + 32)


def test_wave_targets_follow_the_environment():
    """IGM_PAIR_WAVES and IGM_ENERGY_WAVES (read once per process) set the
    row split, and with it the workspace a caller allocates: at 1000 x 1000,
    2048 waves for the two-iteration pass give 17 x 120 waves in 510 blocks,
    6 partial sums each; 4096 for the one-iteration passes 17 x 241 waves in
    1025 blocks, 3 each.  Run in child processes (host code only)."""
    import os
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r); from igm_amd import _lib as L; "
            "print(L.load().igm_energy_workspace_doubles(1, 1000, 1000))" % os.path.dirname(os.path.dirname(
                os.path.abspath(__file__))))

    def run(**env):
        e = {k: v for k, v in os.environ.items() if k not in ("IGM_PAIR_WAVES", "IGM_ENERGY_WAVES")}
        e.update(env)
        return int(subprocess.run([sys.executable, "-c", code], env=e, capture_output=True, text=True,
                                  check=True, timeout=120).stdout.strip())

    assert run() == 1530
    assert run(IGM_PAIR_WAVES="2048") == 6 * 510
    assert run(IGM_ENERGY_WAVES="4096") == 3 * 1025
