"""CPU-side checks of the product library: it loads without a GPU, exports
every entry point the C header declares, and validates arguments before it
touches a device.  Also the host-side API helpers and the SEM generator."""
import math
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT


def header_functions():
    text = open(os.path.join(ROOT, "include", "kpcmi.h")).read()
    return sorted(set(re.findall(r"\b(kpc_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib_path = os.path.join(ROOT, "fastkpc_amd", "libkpcmi.so")
    assert os.path.exists(lib_path), "build the library first (python -m fastkpc_amd.build)"
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (kpc_\w+)", out))
    declared = header_functions()
    assert declared, "no declarations parsed"
    missing = [f for f in declared if f not in exported]
    assert not missing, missing
    from fastkpc_amd import _lib
    L = _lib.lib()
    for f in declared:
        assert hasattr(L, f)


def test_argument_validation_needs_no_gpu():
    from fastkpc_amd import _lib
    from fastkpc_amd._lib import KpcOptions
    import ctypes
    L = _lib.lib()
    o = KpcOptions()
    L.kpc_default_options(o)
    assert o.alpha == 0.1 and o.num_col == 35 and o.max_conditioning_size == -1
    data = np.asfortranarray(np.random.default_rng(0).standard_normal((30, 4)))
    res = ctypes.POINTER(_lib.KpcResult)()
    o.num_col = 40  # >= n
    rc = L.kpc_skeleton_run(_lib.dptr(data), 30, 4, o, ctypes.byref(res))
    assert rc == -1 and b"num_col" in L.kpc_last_error()
    o.num_col = 5
    bad = data.copy()
    bad[3, 1] = np.nan
    rc = L.kpc_skeleton_run(_lib.dptr(np.asfortranarray(bad)), 30, 4, o, ctypes.byref(res))
    assert rc == -1 and b"finite" in L.kpc_last_error()
    o.ci_method = 7
    rc = L.kpc_skeleton_run(_lib.dptr(data), 30, 4, o, ctypes.byref(res))
    assert rc < 0


def test_resolve_max_conditioning_size():
    from fastkpc_amd import resolve_max_conditioning_size as r
    assert r(float("inf"), 48) == 46
    assert r(7, 48) == 7
    assert r(7.9, 48) == 7
    assert r(100, 10) == 8
    with pytest.raises(ValueError):
        r(-1, 10)
    with pytest.raises(ValueError):
        r(float("nan"), 10)


def test_option_whitelist_and_labels():
    import fastkpc_amd as fk
    data = np.zeros((20, 3))
    with pytest.raises(ValueError, match="unknown compatible CUDA option"):
        fk.fastkpc_compatible_cuda_skeleton(data, 0.1, options={"bogus": 1})
    with pytest.raises(ValueError, match="labels length"):
        fk.fastkpc_compatible_cuda_skeleton(data, 0.1, labels=["a"])


def test_sem_generator_is_deterministic_and_acyclic():
    from fastkpc_amd.sem import config_data, generate
    d1, m1, c = config_data("sem351x48")
    d2, m2, _ = config_data("sem351x48")
    assert d1.shape == (351, 48) and np.array_equal(d1, d2) and np.array_equal(m1, m2)
    assert not np.any(np.tril(m1))  # parents precede children
    roots = np.nonzero(~m1.any(axis=0))[0]
    for j in roots:
        assert abs(d1[:, j].std() - 1.0) < 0.15
    d3, _ = generate(n=50, p=6, edge_prob=0.5, seed=11, link="mixed")
    assert np.isfinite(d3).all()


def test_task_seed_matches_oracle_without_gpu():
    # kpc_task_seed is host code: the seed the permutation kernels key on
    import fastkpc_amd as f
    from oracle import orc
    rng = np.random.default_rng(1)
    for method in ("dcc.perm", "hsic.perm"):
        for _ in range(50):
            v = rng.permutation(40)
            S = sorted(int(a) for a in v[2:2 + rng.integers(0, 6)])
            user = int(rng.integers(0, 2 ** 63))
            got = f.task_seed(user, method, int(v[0]), int(v[1]), S)
            assert got == orc.task_seed(user, method, int(v[0]), int(v[1]), S)
            assert got == f.task_seed(user, method, int(v[1]), int(v[0]), S)


def test_regrxons_mixed_validates_before_any_device_call():
    import fastkpc_amd as fk
    data = np.asfortranarray(np.random.default_rng(2).standard_normal((40, 5)))
    with pytest.raises(fk._lib.KpcLibraryError, match=r"\|S\| must be in \[1, 62\]"):
        fk.regrxons_mixed(data, [0, 1], [[2], []])
    with pytest.raises(ValueError, match="one conditioning set per target"):
        fk.regrxons_mixed(data, [0, 1], [[2]])


def test_regression_batch_generator():
    from fastkpc_amd.sem import regression_batch
    d, t, S = regression_batch(fits=64, n=50, seed=9)
    d2, t2, S2 = regression_batch(fits=64, n=50, seed=9)
    assert np.array_equal(d, d2) and np.array_equal(t, t2) and S == S2
    assert d.shape == (50, sum(len(s) + 1 for s in S))
    for tg, s in zip(t, S):
        assert 1 <= len(s) <= 6 and tg == s[-1] + 1 and s == sorted(s)
        X = d[:, s]
        assert X.min() >= -2.0 and X.max() <= 2.0
        noise = d[:, tg] - np.sum(np.sin(2 * X) + 0.3 * X * X, axis=1)
        assert abs(noise.std() - 0.3) < 0.12


def test_plain_c_consumer_compiles_and_runs(tmp_path):
    """tests/c_consumer.c includes kpcmi.h as a C99 program, links
    libkpcmi.so and calls the host-only entry points (defaults, task seeds,
    cache control, argument validation with kpc_last_error)."""
    exe = str(tmp_path / "c_consumer")
    libdir = os.path.join(ROOT, "fastkpc_amd")
    cmd = ["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "c_consumer.c"), "-L", libdir, "-lkpcmi",
           "-Wl,-rpath," + libdir, "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout
