"""Host-side logic of the Python mirror (no device calls)."""
import numpy as np
import pytest


def test_check_defaults_mirror_reference(itcpd):
    c = itcpd.FitCheck(1e-6, 100, 2.5)           # fit_check.jl:16-19
    assert (c.iter, c.counter, c.lastfit, c.final_fit, c.total_iter) == (0, 0, 1.0, 0.0, 0)
    n = itcpd.NoCheck(7)                          # no_check.jl:6
    assert (n.iter, n.max_counter, n.lastfit) == (0, 7, -1.0)
    assert itcpd.CPDFit(c) == 0.0


def test_decompose_solver_type_error_before_any_device_work(itcpd):
    # cp_als.jl:15: a solver that is not a CPDOptimizer is a TypeError
    with pytest.raises(TypeError):
        itcpd.decompose(np.zeros((3, 3, 3)), 4, solver=np.zeros(3))


def test_unknown_algorithm_rejected(itcpd):
    from itensorcpd_jl_amd import api
    assert api._alg_code(None) == 0 and api._alg_code(itcpd.KRPNormal()) == 1
    with pytest.raises(NotImplementedError):
        api._alg_code(object())


def test_target_layout_is_column_major(itcpd):
    from itensorcpd_jl_amd import api
    a = np.arange(24, dtype=np.float32).reshape(2, 3, 4)
    t = api._Target(a)
    assert t.dims == (2, 3, 4) and t.elt == 4 and t.keep.flags.f_contiguous
    t = api._Target(np.arange(6).reshape(2, 3))
    assert t.elt == 8 and t.keep.dtype == np.float64


def test_verbose_trace_format(itcpd, capsys):
    from itensorcpd_jl_amd import api
    c = itcpd.FitCheck(1e-6, 10, 1.0)
    api._print_trace(c, 50, 0, 1.0, np.array([0.6, 0.8]), 2)
    out = capsys.readouterr().out.splitlines()
    assert out[0] == f"50\t 1 \t 0.6 \t {abs(1.0 - 0.6)}"
    assert out[1] == f"50\t 2 \t 0.8 \t {abs(0.6 - 0.8)}"


def test_cpd_container(itcpd):
    f = [np.ones((3, 2)), np.ones((4, 2))]
    cp = itcpd.CPD(f, np.array([1.0, 2.0]))
    assert itcpd.cp_rank(cp) == 2 and cp[1].shape == (4, 2)
    assert cp.copy().isequal(cp)
