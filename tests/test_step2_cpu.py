"""Step-1 hand-off loader (no GPU): numeric arrays load as saved; object
arrays (ragged adjacency) are refused unless unpickling is asked for."""
import numpy as np
import pytest

from tfce_mediation_amd import step2


def test_load_temp_refuses_pickled_arrays_by_default(tmp_path):
    np.save(tmp_path / "num_subjects", 12)
    obj = np.empty(3, dtype=object)
    obj[0], obj[1], obj[2] = [1, 2], [0], [0]
    np.save(tmp_path / "adjac", obj)
    v = step2.load_temp(str(tmp_path), ["num_subjects"])
    assert int(v["num_subjects"]) == 12
    with pytest.raises(ValueError, match="allow_pickle"):
        step2.load_temp(str(tmp_path), ["adjac"])
    v = step2.load_temp(str(tmp_path), ["adjac"], allow_pickle=True)
    assert [list(r) for r in v["adjac"]] == [[1, 2], [0], [0]]
