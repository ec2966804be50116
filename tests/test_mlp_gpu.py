"""The variant-selector MLP's HIP kernels against the CPU oracle."""
import ctypes

import numpy as np
import pytest

import tempi_amd as ta
import typecatalog as tc
from oracle import mlp_ref
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
OUT = ta.lib().tempi_mlp_num_outputs()


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    assert ta.device_count() > 0, "no GPU visible"
    assert ta.lib().tempi_mlp_num_outputs() == OUT


def problem(seed, n, masked):
    rng = np.random.default_rng(seed)
    params = mlp_ref.init_params(OUT, seed)
    params += (rng.standard_normal(params.size) * 0.05).astype(np.float32)
    x = rng.standard_normal((n, 16)).astype(np.float32)
    t = rng.standard_normal((n, OUT)).astype(np.float32)
    m = (rng.random((n, OUT)) > 0.2).astype(np.float32) if masked else None
    return params, x, t, m


@pytest.mark.parametrize("n,masked", [(32, False), (7, True), (1, False), (32, True)])
def test_gradient_matches_oracle(n, masked):
    L = ta.lib()
    params, x, t, m = problem(n + 100 * masked, n, masked)
    grad = np.zeros(params.size, np.float32)
    loss = ctypes.c_float()
    assert L.tempi_mlp_grad(params.ctypes.data, x.ctypes.data, t.ctypes.data,
                            m.ctypes.data if m is not None else None, n, grad.ctypes.data,
                            ctypes.byref(loss)) == 0
    l_ref, g_ref = mlp_ref.loss_and_grad(params, x, t, m, OUT)
    assert abs(loss.value - l_ref) <= 1e-4 * max(1.0, abs(l_ref))
    # fp32 kernel vs float64 oracle
    np.testing.assert_allclose(grad, g_ref, rtol=2e-3, atol=2e-5)


def test_inference_matches_oracle():
    L = ta.lib()
    rng = np.random.default_rng(9)
    params = mlp_ref.init_params(OUT, 9)
    mean = rng.standard_normal(16).astype(np.float32)
    inv = (rng.random(16) + 0.5).astype(np.float32)
    norm = np.concatenate([mean, inv])
    raw = rng.standard_normal((77, 16)).astype(np.float32) * 3
    pred = np.zeros((77, OUT), np.float32)
    best = np.zeros(77, np.int32)
    assert L.tempi_mlp_predict(params.ctypes.data, norm.ctypes.data, raw.ctypes.data, 77,
                               pred.ctypes.data, best.ctypes.data) == 0
    y, _ = mlp_ref.forward(params, (raw - mean) * inv, OUT)
    np.testing.assert_allclose(pred, y, rtol=1e-4, atol=1e-4)
    srt = np.sort(y, axis=1)
    clear = srt[:, 1] - srt[:, 0] > 1e-3
    assert (best[clear] == y.argmin(1)[clear]).all()


def test_one_training_step_is_adam_on_the_oracle_gradient():
    L = ta.lib()
    params, x, t, _ = problem(4, 32, False)
    out = params.copy()
    losses = np.zeros(1, np.float32)
    assert L.tempi_mlp_train(x.ctypes.data, t.ctypes.data, None, 32, 1, ctypes.c_float(1e-3),
                             7, out.ctypes.data, losses.ctypes.data) == 0
    _, g = mlp_ref.loss_and_grad(params, x, t, None, OUT)
    ref, _, _ = mlp_ref.adam_step(params.astype(np.float64), 0 * g, 0 * g, g, 1e-3, 1)
    # a first Adam step moves each parameter by ~lr*sign(g); compare where
    # the gradient is clearly nonzero
    big = np.abs(g) > 1e-4
    np.testing.assert_allclose(out[big], ref[big], rtol=1e-5, atol=2e-6)


def test_training_learns_a_teacher():
    L = ta.lib()
    rng = np.random.default_rng(11)
    teacher = mlp_ref.init_params(OUT, 123)
    x = rng.standard_normal((2048, 16)).astype(np.float32)
    t = mlp_ref.forward(teacher, x, OUT)[0].astype(np.float32)
    params = mlp_ref.init_params(OUT, 1)
    losses = np.zeros(40, np.float32)
    assert L.tempi_mlp_train(x.ctypes.data, t.ctypes.data, None, 2048, 40, ctypes.c_float(2e-3),
                             3, params.ctypes.data, losses.ctypes.data) == 0
    l0, _ = mlp_ref.loss_and_grad(mlp_ref.init_params(OUT, 1), x, t, None, OUT)
    l1, _ = mlp_ref.loss_and_grad(params, x, t, None, OUT)
    assert np.isfinite(losses).all()
    assert l1 < 0.1 * l0, (l0, l1, losses[::8])


def test_commit_time_inference_drives_mpi_pack(mpi, tmp_path):
    """With a model loaded, MPI_Type_commit runs the inference kernel and
    MPI_Pack uses its choice; the packed bytes stay exact whatever it picks."""
    L = ta.lib()
    params = mlp_ref.init_params(OUT, 21)
    norm = np.concatenate([np.zeros(16, np.float32), np.ones(16, np.float32)])
    path = str(tmp_path / "sel.txt")
    assert L.tempi_mlp_save(path.encode(), params.ctypes.data, norm.ctypes.data) == 0
    assert L.tempi_mlp_load(path.encode()) == 0
    try:
        spec = tc.grid_face_subarray(64, 3, "y")
        before = ta.counters()["model_inferences"]
        h = mpi.type_from_spec(spec)
        assert ta.counters()["model_inferences"] == before + 1
        o = orc.OracleType(spec)
        host = orc.splitmix_bytes(o.extent(), 5)
        src = ta.DeviceBuffer(o.extent())
        src.upload(host)
        out = ta.DeviceBuffer(o.size())
        err, pos = mpi.pack(src.addr(), 1, h, out.addr(), o.size())
        assert err == 0
        assert ta.counters()["model_inferences"] == before + 1  # served from the cache
        assert (out.download() == o.pack(host, 0, 1)).all()
        sb, _ = mpi.strided_block(h)
        v_model = ta.select_variant(sb, 1, src.addr(), out.addr(), method=ta.SELECT_MODEL)
        y, _ = mlp_ref.forward(params, np.zeros((1, 16)), OUT)
        assert 0 <= v_model < OUT
        mpi.free(h)
    finally:
        L.tempi_mlp_load(None)


def installed_model():
    """(params, mean, 1/std) of tempi_amd/selector_mlp.txt."""
    import os

    path = os.path.join(os.path.dirname(ta.__file__), "selector_mlp.txt")
    toks = open(path).read().split()
    i = toks.index("norm")
    norm = np.array(toks[i + 1:i + 33], float)
    j = toks.index("params")
    return np.array(toks[j + 1:], float), norm[:16], norm[16:]


def test_commits_in_a_burst_share_inference_launches(mpi):
    """Forty types committed back to back: a commit that finds an inference
    of the library's still running joins the open batch, so there are fewer
    launches than commits.  Every type's two picks then come from the shape
    cache (no inference at lookup) and are, to float rounding, the argmin of
    the CPU reference network on the same features -- a pick stored for the
    wrong type of a batch would not be."""
    L = ta.lib()
    assert L.tempi_mlp_load(None) == 0
    specs = [("vector", 100 + 7 * i, 1 + i % 5, 9 + i, tc.BYTE) for i in range(40)]
    before = ta.counters()["model_inferences"]
    hs = [mpi.type_from_spec(s) for s in specs]
    launches = ta.counters()["model_inferences"] - before
    assert 1 <= launches < len(specs), launches
    params, mean, inv = installed_model()
    src = ta.DeviceBuffer(1 << 20)  # 256-B aligned, as the commit assumed
    dst = ta.DeviceBuffer(1 << 20)
    nf = L.tempi_mlp_num_features()
    f = (ctypes.c_float * nf)()
    picks = []
    for h in hs:
        sb, _ = mpi.strided_block(h)
        for d in (ta.DIR_PACK, ta.DIR_UNPACK):
            picks.append((sb, d, ta.select_variant(sb, 1, src.addr(), dst.addr(), d,
                                                   ta.SELECT_MODEL)))
    after = ta.counters()["model_inferences"]
    # the lookups waited for the batches still running, launched the open
    # one (if there was one), and read every pick from the cache: a lookup
    # the cache missed would have run an inference of its own
    assert after - before - launches <= 1, (launches, after - before)
    for sb, d, v in picks:
        assert L.tempi_mlp_features(ctypes.byref(sb), 1, src.addr(), dst.addr(), d, f) == 0
        x = (np.frombuffer(f, np.float32) - mean) * inv
        y, _ = mlp_ref.forward(params, x[None, :], OUT)
        assert y[0][v] <= y[0].min() + 1e-3, (sb.as_tuple(), d, v, int(y[0].argmin()))
    n_after = ta.counters()["model_inferences"]
    for sb, d, v in picks[:4]:
        assert ta.select_variant(sb, 1, src.addr(), dst.addr(), d, ta.SELECT_MODEL) == v
    assert ta.counters()["model_inferences"] == n_after  # served from the cache
    for h in hs:
        mpi.free(h)


def test_batch_creation_evaluates_new_shapes_in_one_inference():
    """A batch of 200 requests of shapes the selector has not seen is
    created with one inference over all of them (one launch per shape made
    large heterogeneous batches slow to create); each request's variant is,
    to float rounding, the CPU reference network's argmin on its features."""
    L = ta.lib()
    assert L.tempi_mlp_load(None) == 0
    params, mean, inv = installed_model()
    src = ta.DeviceBuffer(64 << 20)
    dst = ta.DeviceBuffer(64 << 20)
    reqs, off = [], 0
    for i in range(200):
        rows, bl = 50 + 13 * i, 1 + (7 * i) % 97
        sb = ta.StridedBlock.make(i % 16, (rows - 1) * (bl + 3 + i) + bl + 16, [bl, rows],
                                  [1, bl + 3 + i])
        assert i % 16 + sb.extent <= src.nbytes and off + sb.size() <= dst.nbytes
        reqs.append((sb, src.addr(), dst.addr(off), 1, ta.DIR_PACK if i % 3 else ta.DIR_UNPACK))
        off += sb.size()
    before = ta.counters()["model_inferences"]
    b = ta.Batch(reqs, method=ta.SELECT_MODEL)
    assert ta.counters()["model_inferences"] - before == 1
    nf = L.tempi_mlp_num_features()
    f = (ctypes.c_float * nf)()
    for i, (sb, s, p, count, d) in enumerate(reqs):
        v = b.variant(i)
        assert L.tempi_mlp_features(ctypes.byref(sb), count, s, p, d, f) == 0
        x = (np.frombuffer(f, np.float32) - mean) * inv
        y, _ = mlp_ref.forward(params, x[None, :], OUT)
        assert y[0][v] <= y[0].min() + 1e-3, (i, v, int(y[0].argmin()))
    b.destroy()


def test_weights_loaded_while_commit_batches_are_pending_replace_their_picks(mpi, tmp_path):
    """Twenty types committed back to back under the installed weights leave
    inferences running and an open batch; loading other weights right after
    drops those picks (the open batch is never launched), so every lookup
    answers with the new network's argmin."""
    L = ta.lib()
    assert L.tempi_mlp_load(None) == 0
    params = mlp_ref.init_params(OUT, 77)
    mean = np.zeros(16, np.float32)
    inv = np.full(16, 0.25, np.float32)
    path = str(tmp_path / "other.txt")
    assert L.tempi_mlp_save(path.encode(), params.ctypes.data,
                            np.concatenate([mean, inv]).ctypes.data) == 0
    specs = [("vector", 300 + 11 * i, 2 + i % 3, 17 + i, tc.BYTE) for i in range(20)]
    hs = [mpi.type_from_spec(s) for s in specs]
    try:
        assert L.tempi_mlp_load(path.encode()) == 0
        src = ta.DeviceBuffer(1 << 20)
        dst = ta.DeviceBuffer(1 << 20)
        nf = L.tempi_mlp_num_features()
        f = (ctypes.c_float * nf)()
        for h in hs:
            sb, _ = mpi.strided_block(h)
            for d in (ta.DIR_PACK, ta.DIR_UNPACK):
                v = ta.select_variant(sb, 1, src.addr(), dst.addr(), d, ta.SELECT_MODEL)
                assert L.tempi_mlp_features(ctypes.byref(sb), 1, src.addr(), dst.addr(), d,
                                            f) == 0
                x = (np.frombuffer(f, np.float32) - mean) * inv
                y, _ = mlp_ref.forward(params, x[None, :], OUT)
                assert y[0][v] <= y[0].min() + 1e-3, (sb.as_tuple(), d, v,
                                                      int(y[0].argmin()))
    finally:
        L.tempi_mlp_load(None)
        for h in hs:
            mpi.free(h)
