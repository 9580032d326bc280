"""continuous_output_batch without a propagation (no GPU): the default-constructed
object's errors and properties (heyoka/test.py:1436-1530)."""

import copy
import pickle

import numpy as np
import pytest

import heyoka_amd as hy

MSG = "Cannot use a default-constructed continuous_output_batch object"


def test_default_constructed():
    for c_out in (hy.continuous_output_batch(), hy.continuous_output_batch_dbl()):
        for arg in (np.zeros((0,)), 1.0, [0.0, 0.0]):
            with pytest.raises(ValueError) as e:
                c_out(arg)
            assert MSG in str(e.value)
        assert c_out.output is None and c_out.times is None and c_out.tcs is None
        for attr in ("bounds", "n_steps"):
            with pytest.raises(ValueError) as e:
                getattr(c_out, attr)
            assert MSG in str(e.value)
        assert c_out.batch_size == 0 and "forward" not in repr(c_out)
        for c2 in (copy.copy(c_out), copy.deepcopy(c_out), pickle.loads(pickle.dumps(c_out))):
            with pytest.raises(ValueError):
                c2.n_steps
            assert c2.batch_size == 0
