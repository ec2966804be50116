"""The shipped variant selector measured against the closed-form rule and
the exhaustive autotuner on copies it was not trained on.  It times every
variant of each copy twice; it is collected last so that a noisy timing
cannot stop the GPU run before the parity tests."""
import numpy as np
import pytest

import tempi_amd as ta

pytestmark = pytest.mark.gpu


def test_installed_model_beats_the_closed_form_rule():
    """The shipped selector, on copies it was not trained on (a seed outside
    the training and held-out ranges), moves more packed bytes per second
    than the closed-form rule and about as many as the exhaustive autotuner.
    Every variant is timed twice, each time the median of three timings: the
    autotuner picks from the first timing, and all three choices are judged
    on the second (the best of one noisy timing, judged on that same timing,
    would be an unreachable bar)."""
    from tempi_amd import selector_train as st

    L = ta.lib()
    assert L.tempi_mlp_load(None) == 0, "tempi_amd/selector_mlp.txt is missing"
    cases = [c for c in st.generate(800, 4242, 8 << 20)
             if (256 << 10) <= c[0].size() * c[4] < st.BIG_MIN]
    cases = cases[:96]
    assert len(cases) >= 64
    src = ta.DeviceBuffer(st.SRC_BYTES)
    src.fill(1)
    dst = ta.DeviceBuffer(st.DST_BYTES + 4096)
    dst.fill(2)
    print(f"timing {len(cases)} copies x {ta.num_variants()} variants, twice", flush=True)
    _, T1, heur, nbytes = st.time_cases(cases, src, dst, log_every=0, rounds=3)
    _, T2, _, _ = st.time_cases(cases, src, dst, log_every=0, rounds=3)
    model = np.array([ta.select_variant(sb, count, src.addr(base), dst.addr(ppos), d,
                                        ta.SELECT_MODEL)
                      for sb, base, d, ppos, count in cases])
    rows = np.arange(len(cases))
    gbs = {
        "model": nbytes.sum() / T2[rows, model].sum() / 1e3,
        "rule": nbytes.sum() / T2[rows, heur].sum() / 1e3,
        "autotuner": nbytes.sum() / T2[rows, T1.argmin(axis=1)].sum() / 1e3,
    }
    print(gbs, flush=True)
    assert gbs["model"] >= 1.05 * gbs["rule"], gbs
    assert gbs["model"] >= 0.95 * gbs["autotuner"], gbs
