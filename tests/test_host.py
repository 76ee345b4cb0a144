"""Host-side checks that need no GPU: the C-ABI library loads and exports every
entry point include/wbx.h declares, and the Python layer raises the
reference's exceptions before any device work."""

import os
import pickle
import re

import numpy as np
import pytest

from wildboar_amd import _lib
from wildboar_amd import distance as wd
from wildboar_amd.variable_len import EOS, get_variable_length, is_end_of_series

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "wbx.h")).read()
    return sorted(set(re.findall(r"\b(wbx_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = header_functions()
    assert "wbx_dev_pairwise" in names and "wbx_pairwise" in names
    for name in names:
        assert hasattr(lib, name), name
    assert lib.wbx_version().decode().startswith("wbx")


def test_python_binding_covers_header():
    for name in header_functions():
        assert name in _lib.EXPORTS, name


def test_metric_validation_matches_reference():
    with pytest.raises(ValueError):
        wd.check_metric("lcss_not_here")
    with pytest.raises(ValueError):
        wd.pairwise_distance(np.zeros((2, 3)), metric="dtw", metric_params={"r": 1.5})
    with pytest.raises(ValueError):
        wd.pairwise_distance(np.zeros((2, 3)), metric="dtw", metric_params={"r": -0.1})
    with pytest.raises(TypeError):
        wd.pairwise_distance(np.zeros((2, 3)), metric="dtw", metric_params={"window": 0.1})
    with pytest.raises(TypeError):
        wd.pairwise_distance(np.zeros((2, 3)), metric="euclidean", metric_params={"r": 0.1})
    with pytest.raises(ValueError):
        wd.pairwise_distance(np.zeros((2, 3)), metric="wdtw", metric_params={"g": -1.0})
    m = wd.check_metric("wdtw")(r=0.3, g=0.2)
    m2 = pickle.loads(pickle.dumps(m))
    assert (m2.name, m2.r, m2.g) == ("wdtw", 0.3, 0.2)


def test_input_validation_matches_reference():
    x = np.zeros((3, 5))
    bad = x.copy()
    bad[1, 2] = np.nan
    with pytest.raises(ValueError, match="NaN"):
        wd.pairwise_distance(bad, x)
    bad[1, 2] = np.inf
    with pytest.raises(ValueError, match="infinity"):
        wd.pairwise_distance(bad, x)
    bad[1, 2] = EOS
    with pytest.raises(ValueError, match="equal length"):
        wd.pairwise_distance(bad, x)
    with pytest.raises(ValueError, match="non-elastic"):
        wd.pairwise_distance(np.zeros((2, 4)), np.zeros((2, 5)), metric="euclidean")
    with pytest.raises(ValueError):
        wd.pairwise_distance(np.zeros((2, 4)), np.zeros((2, 4)), dim=3)
    with pytest.raises(ValueError):
        wd.argmin_distance(np.zeros((2, 4)), np.zeros((3, 4)), k=0)
    with pytest.raises(ValueError):
        wd.argmin_distance(np.zeros((2, 4)), np.zeros((3, 4)), dim=1)
    with pytest.raises(ValueError):
        wd.argmin_distance(np.zeros((2, 4)), np.zeros((3, 4)), lower_bound=np.zeros((3, 2)))
    with pytest.raises(ValueError, match="cannot be empty"):
        wd.pairwise_subsequence_distance([], np.zeros((2, 4)))
    with pytest.raises(ValueError, match="Invalid subsequnce shape"):
        wd.pairwise_subsequence_distance([np.zeros(6)], np.zeros((2, 4)))
    with pytest.raises(ValueError):
        wd.paired_distance(np.zeros((2, 4)), np.zeros((3, 4)))
    assert wd.pairwise_distance(np.zeros(5)) == 0.0


def test_eos_marker_layout():
    assert np.isnan(EOS)
    bits32 = np.array(EOS, dtype=np.float32).view(np.uint32)
    assert int(bits32) & 0xF == 9
    bits64 = np.array(EOS, dtype=np.float64).view(np.uint64)
    assert int(bits64) == 0x7FF8000120000000
    x = np.array([[1.0, 2.0, EOS, EOS], [1.0, 2.0, 3.0, 4.0], [EOS, EOS, EOS, EOS]])
    assert get_variable_length(x).tolist() == [2, 4, 0]
    assert is_end_of_series(np.array([np.nan, EOS])).tolist() == [False, True]
    assert get_variable_length(np.array([1.0, EOS])) == 1
