"""The allocation-free current_green (qclhip_current_green_noalloc) against
qclhip_current_green on the same inputs: two calls in a row, the second with
a smaller problem, each equal bit for bit (same kernels; only the storage of
the block traces differs), and the size limit of its fixed scratch."""
import numpy as np
import pytest

import scba_fixture as fx

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    import torch
    assert torch.cuda.is_available(), "GPU tests need the MI355X"
    from qclhip import api
    api.context(0)
    return api


def _case(hip, N, NE, Nk, seed):
    w = fx.syn.gf_workload(N=N, NE=NE, Nk=Nk, seed=seed)
    g = hip.Grid(w["eps"], w["wE"], w["wk"], N)
    fams = [fx.syn.diag_fields_device(fx.syn.injection_family(w)), fx.syn.diag_fields_device(w)]
    return g, hip.h_to_device(w["h"]), fams


def test_noalloc_equals_current_green(hip):
    import torch
    for N, NE, Nk, seed in ((24, 512, 16, 3), (7, 45, 3, 4)):  # the second call is smaller
        print(f"noalloc case N={N} NE={NE} Nk={Nk}", flush=True)
        g, h, fams = _case(hip, N, NE, Nk, seed)
        ref, ri = hip.current_green_production(g, h, fams, 1.0)
        target = 0.9 * ri["number"]
        ref, ri = hip.current_green_production(g, h, fams, target)
        got, gi = hip.current_green_production(g, h, fams, target, allocation_free=True)
        for k in ("GR", "GL", "GG", "A", "scale"):
            assert torch.equal(getattr(got, k), getattr(ref, k)), k
        assert gi == ri


def test_noalloc_size_limit(hip):
    from qclhip import ArgumentError
    g, h, fams = _case(hip, 1, 4097, 128, 5)  # 524,416 blocks: one field is 8 MiB
    with pytest.raises(ArgumentError, match="at most NE"):
        hip.current_green_production(g, h, fams, 1.0, allocation_free=True)
