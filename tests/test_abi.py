"""The C ABI library loads and exports every symbol include/ptg.h declares
(no GPU needed: no compute call is made)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "ptg.h")).read()
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(ptg_\w+)\s*\(", text, re.M)))


def test_header_lists_entry_points():
    syms = declared_symbols()
    for s in ("ptg_create", "ptg_run", "ptg_collide", "ptg_fetch", "ptg_push", "ptg_histogram", "ptg_eval"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    from ptarmigan_amd import _abi
    lib = ctypes.CDLL(_abi.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    assert set(_abi.EXPORTS) == set(declared_symbols())


def test_struct_sizes_match_header():
    """ctypes mirrors of the ABI structs have the sizes the C compiler gives them."""
    from ptarmigan_amd import _abi
    src = os.path.join(ROOT, "include", "ptg.h")
    code = "#include <stdio.h>\n#include \"%s\"\nint main(){printf(\"%%zu %%zu %%zu %%zu %%zu %%zu %%zu %%zu %%zu %%zu %%zu %%d\\n\"," \
           "sizeof(ptg_laser_desc),sizeof(ptg_collide_options),sizeof(ptg_particles),sizeof(ptg_counters)," \
           "sizeof(ptg_beam_desc),sizeof(ptg_hist_spec),sizeof(ptg_surrogate_model),sizeof(ptg_surrogate_train_desc)," \
           "sizeof(ptg_surrogate_report),sizeof(ptg_tune_result),sizeof(ptg_stats_spec),PTG_MLP_NPARAM);}\n" % src
    import subprocess
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "s.c")
        open(c, "w").write(code)
        subprocess.run(["gcc", "-o", os.path.join(d, "s"), c], check=True)
        out = subprocess.run([os.path.join(d, "s")], capture_output=True, text=True, check=True).stdout.split()
    want = [ctypes.sizeof(t) for t in (_abi.LaserDesc, _abi.CollideOptions, _abi.Particles, _abi.Counters,
                                       _abi.BeamDesc, _abi.HistSpec, _abi.SurrogateModel, _abi.SurrogateTrainDesc,
                                       _abi.SurrogateReport, _abi.TuneResult, _abi.StatsSpec)]
    assert [int(x) for x in out[:11]] == want
    assert int(out[11]) == _abi.MLP_NPARAM


def test_engine_without_gpu_fails_loudly():
    """On a host without a GPU the engine raises instead of falling back."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from ptarmigan_amd import Engine, PtgError
    with pytest.raises(PtgError):
        Engine(0)
