"""Training options on the host (no GPU): the C ABI's defaults are the
reference's TrainOptions defaults (MParT/TrainMap.h:19-40), and the Python
mirror rejects optimizers it does not provide before any device call."""
import ctypes
import math

import pytest

import mpart_amd as mt


def test_default_train_options_match_reference():
    o = mt._CTrainOptions()
    mt._lib.mpart_default_train_options(ctypes.byref(o))
    assert o.algorithm == 0
    assert o.stopval == -math.inf and o.maxtime == math.inf
    assert (o.ftol_rel, o.ftol_abs, o.xtol_rel, o.xtol_abs) == (1e-3, 1e-3, 1e-4, 1e-4)
    assert o.maxeval == 1000 and o.verbose == 0 and o.history == 10


def test_python_train_options():
    t = mt.TrainOptions()
    assert t.opt_alg == "LD_SLSQP" and t.opt_maxeval == 1000 and t.opt_ftol_rel == 1e-3
    c = mt.TrainOptions(opt_alg="LD_LBFGS", opt_maxeval=7, history=4)._c()
    # the reference default LD_SLSQP runs L-BFGS here, and says so
    with pytest.warns(RuntimeWarning, match="L-BFGS"):
        mt.TrainOptions()._c()
    assert c.maxeval == 7 and c.history == 4 and c.xtol_abs == 1e-4
    with pytest.raises(NotImplementedError):
        mt.TrainOptions(opt_alg="LN_NELDERMEAD")._c()
    with pytest.raises(AttributeError):
        mt.TrainOptions(opt_tolerance=1.0)
    assert "opt_maxeval = 1000" in str(mt.TrainOptions())
