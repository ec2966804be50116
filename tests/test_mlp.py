"""The variant-selector MLP on the CPU: the oracle's gradient against finite
differences (and a planted mistake against the same check), the feature
extractor against its definition, and the model file format."""
import ctypes
import math
import os

import numpy as np
import pytest

import tempi_amd as ta
from oracle import mlp_ref

OUT = ta.lib().tempi_mlp_num_outputs()


def fd_grad(params, x, t, m, idx, h=1e-6):
    g = []
    for i in idx:
        p1 = params.astype(np.float64).copy()
        p2 = p1.copy()
        p1[i] += h
        p2[i] -= h
        l1, _ = mlp_ref.loss_and_grad(p1, x, t, m, OUT)
        l2, _ = mlp_ref.loss_and_grad(p2, x, t, m, OUT)
        g.append((l1 - l2) / (2 * h))
    return np.array(g)


def sample_problem(seed, n=9):
    rng = np.random.default_rng(seed)
    params = mlp_ref.init_params(OUT, seed).astype(np.float64)
    params += rng.standard_normal(params.size) * 0.05  # nonzero biases
    x = rng.standard_normal((n, 16))
    t = rng.standard_normal((n, OUT))
    m = (rng.random((n, OUT)) > 0.2).astype(np.float64)
    return params, x, t, m


def test_oracle_gradient_matches_finite_differences():
    params, x, t, m = sample_problem(1)
    _, g = mlp_ref.loss_and_grad(params, x, t, m, OUT)
    rng = np.random.default_rng(2)
    n = params.size
    # every parameter block: W1, b1, W2, b2, W3, b3
    offs = np.cumsum([0] + [int(np.prod(s)) for s in mlp_ref.shapes(OUT)])
    idx = np.concatenate([rng.integers(offs[k], offs[k + 1], 12) for k in range(6)])
    assert offs[-1] == n == ta.lib().tempi_mlp_num_params()
    fd = fd_grad(params, x, t, m, idx)
    np.testing.assert_allclose(g[idx], fd, rtol=1e-5, atol=1e-8)


def test_planted_gradient_mistake_is_caught():
    """A backward pass that forgets the ReLU mask of the second layer must
    fail the finite-difference check."""
    params, x, t, m = sample_problem(3)

    def wrong(params):
        W1, b1, W2, b2, W3, b3 = mlp_ref.unpack(params, OUT)
        y, (xx, h1, h2) = mlp_ref.forward(params, x, OUT)
        dy = 2.0 * m * (y - t) / m.sum()
        d2 = dy @ W3.T  # mask dropped
        return h1.T @ d2

    offs = np.cumsum([0] + [int(np.prod(s)) for s in mlp_ref.shapes(OUT)])
    idx = np.arange(offs[2], offs[3])
    fd = fd_grad(params, x, t, m, idx).reshape(64, 64)
    assert not np.allclose(wrong(params), fd, rtol=1e-4, atol=1e-7)


def test_adam_step_closed_form():
    # first step of Adam moves every parameter by lr * sign(g) (bias-corrected)
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -4.0, 1e-3])
    p2, m1, m2 = mlp_ref.adam_step(p, np.zeros(3), np.zeros(3), g, 0.01, 1)
    np.testing.assert_allclose(p - p2, 0.01 * np.sign(g), rtol=1e-4)


def test_features_follow_their_definition():
    L = ta.lib()
    f = (ctypes.c_float * 16)()
    # x face of the float grid at its real addresses' alignment
    sb = ta.StridedBlock.make(3226116, 555967328, [12, 512, 512], [1, 2072, 1073296])
    assert L.tempi_mlp_features(ctypes.byref(sb), 1, 1 << 20, 1 << 21, 0, f) == 0
    exp = [math.log2(12), math.log2(512 * 512), math.log2(12 * 512 * 512), 2,
           math.log2(512), math.log2(2072), math.log2(512), math.log2(1073296), 0,
           2,  # 3226116 = 4 * odd
           4, 3,  # strides 2072 = 8 * 259
           2, 0, 1, math.log2(2072 / 12)]
    np.testing.assert_allclose(list(f), exp, rtol=1e-5)
    # contiguous type: one row, unpack
    sb = ta.StridedBlock.make(0, 4096, [4096], [1])
    assert L.tempi_mlp_features(ctypes.byref(sb), 3, 1 << 20, (1 << 20) + 3, 1, f) == 0
    exp = [math.log2(3 * 4096), 0, math.log2(3 * 4096), 0, 0, 0, 0, 0, 0, 4, 0, 4, 4, 1, 0, 0]
    np.testing.assert_allclose(list(f), exp, rtol=1e-5)


def test_model_file_round_trip_and_rejects(tmp_path):
    L = ta.lib()
    n = L.tempi_mlp_num_params()
    params = mlp_ref.init_params(OUT, 5)
    norm = np.concatenate([np.arange(16, dtype=np.float32), np.ones(16, np.float32)])
    path = str(tmp_path / "m.txt")
    assert L.tempi_mlp_save(path.encode(), params.ctypes.data, norm.ctypes.data) == 0
    text = open(path).read().split()
    assert text[:2] == ["tempi-mlp", "1"]
    vals = np.array([float(v) for v in text[text.index("params") + 1:]], np.float32)
    assert vals.size == n and (vals == params).all()
    assert L.tempi_mlp_load(path.encode()) == 0 and L.tempi_mlp_loaded() == 1
    bad = tmp_path / "bad.txt"
    bad.write_text(open(path).read().replace(f"out {OUT}", f"out {OUT - 1}"))
    assert L.tempi_mlp_load(str(bad).encode()) == ta.TEMPI_ERR_IO
    assert L.tempi_mlp_loaded() == 0
    trunc = tmp_path / "trunc.txt"
    trunc.write_text(open(path).read()[:2000])
    assert L.tempi_mlp_load(str(trunc).encode()) == ta.TEMPI_ERR_IO
    assert L.tempi_mlp_load(b"/nonexistent/model.txt") == ta.TEMPI_ERR_IO
    extra = tmp_path / "extra.txt"
    extra.write_text(open(path).read() + "0.5\n")
    assert L.tempi_mlp_load(str(extra).encode()) == ta.TEMPI_ERR_IO
    # seeded mutations: each file loads whole or is refused
    good = open(path).read()
    rng = np.random.default_rng(12)
    refused = 0
    for k in range(200):
        t = list(good)
        i = int(rng.integers(0, len(t)))
        t[i] = "x-. 9e,+"[int(rng.integers(0, 8))]
        if k % 3 == 0:
            del t[int(rng.integers(0, len(t))):]
        mut = tmp_path / "mut.txt"
        mut.write_text("".join(t))
        err = L.tempi_mlp_load(str(mut).encode())
        assert err in (0, ta.TEMPI_ERR_IO), k
        assert L.tempi_mlp_loaded() == (1 if err == 0 else 0), k
        refused += err != 0
    assert 0 < refused < 200
    # the installed weights files load
    for name, load in (("selector_mlp.txt", L.tempi_mlp_load),
                       ("method_mlp.txt", L.tempi_method_model_load)):
        installed = os.path.join(os.path.dirname(ta.LIB_PATH), name)
        if os.path.exists(installed):
            assert load(installed.encode()) == 0, name
    # without a model the selection is the closed-form rule
    sb = ta.StridedBlock.make(0, 524032, [256, 1024], [1, 512])
    assert ta.select_variant(sb, 1, 1 << 20, 1 << 20, method=ta.SELECT_MODEL) == \
        ta.select_variant(sb, 1, 1 << 20, 1 << 20, method=ta.SELECT_HEURISTIC)
    L.tempi_mlp_load(None)  # back to the default model, if one is installed


def test_method_features_follow_their_definition():
    """The staging-method network's inputs: log2 bytes, row bytes, rows,
    stride and stride / row bytes; same node; log2 alignment of the row
    length (capped at 16 B); zeros after."""
    L = ta.lib()
    f = (ctypes.c_float * 16)()
    assert L.tempi_method_features(4096, 64, 192, 1, f) == 0
    exp = [12, 6, 6, math.log2(192), math.log2(3), 1, 4] + [0] * 9
    np.testing.assert_allclose(list(f), exp, rtol=1e-6)
    # 12-B rows of a float grid's x face, another node
    assert L.tempi_method_features(3 << 20, 12, 2072, 0, f) == 0
    exp = [math.log2(3 << 20), math.log2(12), math.log2((3 << 20) / 12), math.log2(2072),
           math.log2(2072 / 12), 0, 2] + [0] * 9
    np.testing.assert_allclose(list(f), exp, rtol=1e-6)
    assert L.tempi_method_features(-1, 12, 2072, 0, f) == ta.TEMPI_ERR_ARG


def test_method_model_file_is_checked_like_the_selector(tmp_path):
    """The method model is the selector's network with its own weights: a
    file of the wrong shape is refused and the library keeps the tables."""
    L = ta.lib()
    params = mlp_ref.init_params(OUT, 9)
    norm = np.concatenate([np.zeros(16, np.float32), np.ones(16, np.float32)])
    path = str(tmp_path / "method.txt")
    assert L.tempi_mlp_save(path.encode(), params.ctypes.data, norm.ctypes.data) == 0
    bad = tmp_path / "bad.txt"
    bad.write_text(open(path).read().replace(f"out {OUT}", f"out {OUT - 1}"))
    assert L.tempi_method_model_load(str(bad).encode()) == ta.TEMPI_ERR_IO
    assert L.tempi_method_model_loaded() == 0
    assert L.tempi_method_model_load(path.encode()) == 0
    assert L.tempi_method_model_loaded() == 1
    assert L.tempi_method_model_unload() == 0 and L.tempi_method_model_loaded() == 0
    L.tempi_method_model_load(None)  # back to the installed model, if any
