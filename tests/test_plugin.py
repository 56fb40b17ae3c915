"""The drop-in plugin lib/libpostprocessor.so, driven by a stand-in host
(tests/plugin/plugin_host.cpp) through the extern "C" exports and the
Interpolator / Data vtables, as the reference's host application does.

CPU: exports, config parsing and abort paths that need no device.
GPU: the reference's own known-answer test (LinearBasis/tests/InterpolateArray/test.cpp:
surplus.plt at x_d = (d+1)/118, 118 values within EPSILON = 1e-3) through the
plugin, plus parity with the oracle to 1e-12, the multistate overload and the
device-side abort paths.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import hddm_lb
import oracle
from conftest import GOLDEN, ROOT

HOST = os.path.join(ROOT, "tests", "plugin", "_build", "plugin_host")
PLUGIN = os.path.join(ROOT, "hddm-solver-postprocessors_amd", "lib", "libpostprocessor.so")
EXPORTS = ["getInterpolator", "getData", "isSupported", "getPriority", "resetInterpolator",
           "tryAcquireDevice", "releaseDevice", "getDeviceProperties"]


def run_host(tmp_path, mode, *args, cfg="n_agents 59\npriorityHIP 4\n", timeout=120):
    if not os.path.exists(HOST) or not os.path.exists(PLUGIN):
        pytest.fail("plugin or stand-in host not built (run __graft_entry__.build())")
    if cfg is not None:
        (tmp_path / "hddm-solver.cfg").write_text(cfg)
    shutil.copy(os.path.join(GOLDEN, "surplus.plt"), tmp_path / "surplus.plt")
    p = subprocess.run([HOST, PLUGIN, mode] + [str(a) for a in args], cwd=tmp_path,
                       capture_output=True, text=True, timeout=timeout)
    print(p.stdout[-2000:], p.stderr[-2000:], flush=True)
    return p


def parse_value_line(out):
    for line in out.splitlines():
        f = line.split()
        if f and f[0] == "value":
            return np.array([float(v) for v in f[1:]])
    raise AssertionError("no value line in the host output")


def test_plugin_exports_every_entry_point(tmp_path):
    p = run_host(tmp_path, "exports")
    assert p.returncode == 0
    for name in EXPORTS:
        assert "export %s found" % name in p.stdout


def test_plugin_reads_priority_from_cfg(tmp_path):
    p = run_host(tmp_path, "kat", "surplus.plt", cfg="! comment line\nn_agents 59\npriorityHIP 7 ! GPU priority\n")
    assert "priority 7" in p.stdout
    if hddm_lb.device_count() == 0:
        assert p.returncode == 2 and "no device" in p.stderr


def test_plugin_missing_cfg_aborts(tmp_path):
    p = run_host(tmp_path, "kat", "surplus.plt", cfg=None)
    assert p.returncode == 3
    assert "Cannot open config file" in p.stderr


def test_plugin_missing_required_key_aborts(tmp_path):
    p = run_host(tmp_path, "kat", "surplus.plt", cfg="n_agents 59\n")
    assert p.returncode == 3
    assert 'Required parameter "priority" is undefined' in p.stderr


def test_plugin_bad_precision_aborts(tmp_path):
    p = run_host(tmp_path, "kat", "surplus.plt", cfg="n_agents 59\npriorityHIP 1\nprecisionHIP fp8\n")
    assert p.returncode == 3
    assert "precisionHIP" in p.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("f16x2", 1e-6)])
def test_plugin_known_answer(tmp_path, kat, precision, tol):
    cfg = "n_agents 59\npriorityHIP 4\nprecisionHIP %s\n" % precision
    p = run_host(tmp_path, "kat", "surplus.plt", 3, cfg=cfg)
    assert p.returncode == 0
    assert "supported 1" in p.stdout and "second acquire refused" in p.stdout
    y = parse_value_line(p.stdout)
    np.testing.assert_allclose(y, kat["expected"], rtol=0, atol=kat["epsilon"])
    st = kat["state"]
    yo = oracle.forward(st["index_i"], st["index_j"], st["surplus"], np.array([kat["x"]]))[0]
    assert np.abs(y - yo).max() / np.abs(yo).max() <= tol


@pytest.mark.gpu
def test_plugin_multistate_overload(tmp_path, kat):
    p = run_host(tmp_path, "multistate", "surplus.plt", 5)
    assert p.returncode == 0
    st = kat["state"]
    xs, ys = {}, {}
    for line in p.stdout.splitlines():
        f = line.split()
        if f and f[0] in ("x", "value") and len(f) > 2:
            (xs if f[0] == "x" else ys)[int(f[1])] = np.array([float(v) for v in f[2:]])
    assert sorted(ys) == list(range(5))
    X = np.array([xs[s] for s in range(5)])
    yo = oracle.forward(st["index_i"], st["index_j"], st["surplus"], X)
    Y = np.array([ys[s] for s in range(5)])
    assert np.abs(Y - yo).max() / np.abs(yo).max() <= 1e-12


@pytest.mark.gpu
def test_plugin_reload_aborts(tmp_path):
    p = run_host(tmp_path, "reload", "surplus.plt")
    assert p.returncode == 3
    assert "State 0 data is already loaded" in p.stderr


@pytest.mark.gpu
def test_plugin_dof_mismatch_aborts(tmp_path):
    p = run_host(tmp_path, "dofs", "surplus.plt")
    assert p.returncode == 3
    assert "mismatches actual data TotalDof" in p.stderr
